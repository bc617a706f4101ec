// case.cpp -- TRITON input formats and the ghost-padded case setup.
// Each routine cites the reference behaviour it keeps; parsing quirks that
// change values (atoi for tokens without '.', first key wins, '#' removal of
// one character) are kept on purpose so inputs load to the same numbers.
#include "case.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tr {
namespace {

std::string trim(std::string s) {
  auto notsp = [](unsigned char ch) { return !std::isspace(ch); };
  s.erase(s.begin(), std::find_if(s.begin(), s.end(), notsp));
  s.erase(std::find_if(s.rbegin(), s.rend(), notsp).base(), s.end());
  return s;
}

// std::getline-based split: empty fields are kept except a trailing one.
std::vector<std::string> split(const std::string& s, char delim) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, delim)) out.push_back(item);
  return out;
}

std::string unquote(std::string v) {
  if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
  else if (v.size() == 1 && v == "\"") v.clear();
  return v;
}

// istringstream >> double semantics (src/string_utils.h vecstr_to_vecflt).
double to_double(const std::string& s) {
  std::istringstream iss(s);
  double v = 0.0;
  iss >> v;
  return v;
}
int to_int(const std::string& s) {
  std::istringstream iss(s);
  int v = 0;
  iss >> v;
  return v;
}

std::string resolve(const std::string& p, const std::string& base) {
  if (p.empty() || p[0] == '/' || base.empty()) return p;
  return base + "/" + p;
}

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) throw InputError("Error reading file: " + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string get(const std::map<std::string, std::string>& m, const std::string& k,
                const std::string& dflt = "") {
  auto it = m.find(k);
  return it == m.end() ? dflt : it->second;
}

// src/matrix_io.h:121-130 -- '.' selects atof, anything else atoi.
inline double token_value(const std::string& tok) {
  if (tok.find('.') != std::string::npos) return std::atof(tok.c_str());
  return double(std::atoi(tok.c_str()));
}

// x/y CSV of source or observation points (src/config_utils.h:241-302).
void parse_points(const std::string& path, std::vector<double>& xs, std::vector<double>& ys) {
  std::ifstream f(path);
  if (!f.good()) throw InputError("Error reading file: " + path);
  std::string line;
  while (std::getline(f, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '%') continue;
    size_t hash = line.find('#');
    if (hash != std::string::npos && hash > 0 && line[hash - 1] != '\'' && line[hash - 1] != '"')
      line.erase(line.begin() + long(hash));
    auto kv = split(line, ',');
    if (kv.size() < 2) throw InputError("Malformed point line in " + path + ": " + line);
    xs.push_back(to_double(unquote(kv[0])));
    ys.push_back(to_double(unquote(kv[1])));
  }
}

// src/triton_init.h:1595-1605
int calc_col(double x, double xll, double cs) { return int(std::ceil(((x - xll) / cs) + 1e-16) - 1); }
int calc_row(double y, double yll, double cs, int nrows) {
  return int(std::ceil((nrows - ((y - yll) / cs)) + 1e-16) - 1);
}

std::vector<double> pad(const std::vector<double>& in, int nr, int nc, double ghost) {
  std::vector<double> out(size_t(nr + 2) * size_t(nc + 2), ghost);
  for (int i = 0; i < nr; i++)
    std::copy(in.begin() + long(i) * nc, in.begin() + long(i + 1) * nc,
              out.begin() + long(i + 1) * (nc + 2) + 1);
  return out;
}

// src/matrix.h:848-865 -- ghosts take the nearest inner value (rows first,
// then columns, so corners take the diagonal inner value).
template <class T>
void copy_into_ghosts(std::vector<T>& a, int rows, int cols) {
  for (int i = 0; i < rows; i++) {
    a[size_t(i) * cols] = a[size_t(i) * cols + 1];
    a[size_t(i) * cols + cols - 1] = a[size_t(i) * cols + cols - 2];
  }
  for (int j = 0; j < cols; j++) {
    a[j] = a[size_t(cols) + j];
    a[size_t(rows - 1) * cols + j] = a[size_t(rows - 2) * cols + j];
  }
}

// src/matrix.h:929-945
void set_walls(std::vector<double>& z, int rows, int cols) {
  for (int i = 0; i < rows; i++) {
    z[size_t(i) * cols] = 1e6;
    z[size_t(i) * cols + cols - 1] = 1e6;
  }
  for (int j = 0; j < cols; j++) {
    z[j] = 1e6;
    z[size_t(rows - 1) * cols + j] = 1e6;
  }
}

std::vector<double> load_raster(const Config& cfg, const std::string& path, int nr, int nc) {
  if (cfg.input_format == "ASC") return read_ascii_raster(path, nr, nc, 0, true);
  return read_binary_raster(path, nr, nc, 2, true);
}

}  // namespace

// src/config_utils.h:195-238
std::map<std::string, std::string> parse_cfg_text(const std::string& text) {
  std::map<std::string, std::string> out;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    size_t hash = line.find('#');
    if (hash != std::string::npos && line[hash - 1] != '\'' && line[hash - 1] != '"')
      line.erase(line.begin() + long(hash));
    auto kv = split(line, '=');
    if (kv.empty()) continue;
    std::string key = kv[0];
    std::string value = kv.size() > 1 ? kv[1] : std::string();
    if (!value.empty() && value.front() == '"' && value.back() == '"')
      value = value.size() >= 2 ? value.substr(1, value.size() - 2) : std::string();
    for (size_t i = 2; i < kv.size(); i++) value += kv[i];
    out.insert({key, value});  // first occurrence wins, as std::map::insert
  }
  return out;
}

// src/config_utils.h:422-517
Config config_from_text(const std::string& text, const std::string& base_dir) {
  Config c;
  c.raw = parse_cfg_text(text);
  const auto& m = c.raw;
  c.outfile_pattern = get(m, "outfile_pattern");
  c.hydrograph_filename = resolve(get(m, "hydrograph_filename"), base_dir);
  c.runoff_filename = resolve(get(m, "runoff_filename"), base_dir);
  c.dem_filename = resolve(get(m, "dem_filename"), base_dir);
  c.header_filename = resolve(get(m, "header_filename"), base_dir);
  c.src_loc_file = resolve(get(m, "src_loc_file"), base_dir);
  c.runoff_map = resolve(get(m, "runoff_map"), base_dir);
  c.observation_loc_file = resolve(get(m, "observation_loc_file"), base_dir);
  c.extbc_file = resolve(get(m, "extbc_file"), base_dir);
  c.extbc_dir = resolve(get(m, "extbc_dir"), base_dir);
  c.const_mann = std::atof(get(m, "const_mann").c_str());
  c.time_step = std::atof(get(m, "time_step").c_str());
  c.time_increment_fixed = std::atoi(get(m, "time_increment_fixed").c_str());
  c.time_series_flag = std::atoi(get(m, "time_series_flag").c_str());
  c.courant = std::atof(get(m, "courant").c_str());
  c.hextra = std::atof(get(m, "hextra").c_str());
  c.open_boundaries = std::atoi(get(m, "open_boundaries").c_str());
  c.num_sources = std::atoi(get(m, "num_sources").c_str());
  c.num_runoffs = std::atoi(get(m, "num_runoffs").c_str());
  c.num_extbc = std::atoi(get(m, "num_extbc").c_str());
  c.checkpoint_id = std::atoi(get(m, "checkpoint_id", "0").c_str());
  c.it_count = std::atoi(get(m, "it_count", "0").c_str());
  c.it_print = std::atoi(get(m, "it_print", "10000").c_str());
  c.print_option = get(m, "print_option");
  std::transform(c.print_option.begin(), c.print_option.end(), c.print_option.begin(),
                 [](unsigned char ch) { return char(std::tolower(ch)); });
  c.max_value_print_option = get(m, "max_value_print_option");
  c.input_format = get(m, "input_format");
  c.input_option = get(m, "input_option");
  if (c.input_option.empty()) c.input_option = "SEQ";
  c.output_format = get(m, "output_format");
  c.output_option = get(m, "output_option");
  c.h_infile = resolve(get(m, "h_infile"), base_dir);
  c.qx_infile = resolve(get(m, "qx_infile"), base_dir);
  c.qy_infile = resolve(get(m, "qy_infile"), base_dir);
  c.n_infile = resolve(get(m, "n_infile"), base_dir);
  c.sim_start_time = std::atof(get(m, "sim_start_time").c_str());
  c.sim_duration = std::atof(get(m, "sim_duration").c_str());
  std::string pi = get(m, "print_interval");
  c.print_observation = std::atof(get(m, "print_observation", pi).c_str());
  c.print_interval = std::atof(pi.c_str());
  c.output_folder = get(m, "output_folder", "output");

  if (c.num_sources > 0) parse_points(c.src_loc_file, c.src_x, c.src_y);
  if (c.time_series_flag) parse_points(c.observation_loc_file, c.obs_x, c.obs_y);
  if (c.num_extbc > 0) {
    // src/config_utils.h:305-419
    std::ifstream f(c.extbc_file);
    if (!f.good()) throw InputError("Error reading file: " + c.extbc_file);
    std::string line;
    while (std::getline(f, line)) {
      line = trim(line);
      if (line.empty() || line[0] == '%') continue;
      size_t hash = line.find('#');
      if (hash != std::string::npos && hash > 0 && line[hash - 1] != '\'' && line[hash - 1] != '"')
        line.erase(line.begin() + long(hash));
      auto kv = split(line, ',');
      if (kv.size() < 5) throw InputError("Malformed extbc line: " + line);
      int type = to_int(unquote(kv[0]));
      std::string fname = "0.0";
      if (type != 0) {
        if (kv.size() < 6) throw InputError("Missing extbc value: " + line);
        fname = unquote(kv[5]);
      }
      c.extbc_type.push_back(type);
      c.extbc_x1.push_back(to_double(unquote(kv[1])));
      c.extbc_y1.push_back(to_double(unquote(kv[2])));
      c.extbc_x2.push_back(to_double(unquote(kv[3])));
      c.extbc_y2.push_back(to_double(unquote(kv[4])));
      c.extbc_fname.push_back(type == 1 ? c.extbc_dir + "/" + fname : fname);
    }
    if (int(c.extbc_type.size()) < c.num_extbc)
      throw InputError("extbc file lists fewer segments than num_extbc");
  }
  return c;
}

Config load_config(const std::string& cfg_path, const std::string& base_dir) {
  return config_from_text(read_file(cfg_path), base_dir);
}

// src/inflow.h:176-219 and src/extbc.h:142-165 -- CSV rows, '%' comments.
Table load_table_csv(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw InputError("Error reading file: " + path);
  Table t;
  std::string line;
  while (std::getline(f, line)) {
    line = trim(line);
    auto parts = split(line, '%');
    std::string body = parts.empty() ? std::string() : trim(parts[0]);
    if (body.empty()) continue;
    std::vector<double> row;
    for (auto& tok : split(body, ',')) row.push_back(to_double(tok));
    t.rows.push_back(row);
  }
  return t;
}

// src/dem_utils.h:164-233 -- six lines, the last token of each, by position.
DemHeader read_dem_header_ascii(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw InputError("Error reading file: " + path);
  DemHeader hd;
  std::string line;
  for (int ln = 1; ln <= 6 && std::getline(f, line); ln++) {
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    auto toks = split(line, ' ');
    const char* v = toks.empty() ? "" : toks.back().c_str();
    switch (ln) {
      case 1: hd.ncols = std::atoi(v); break;
      case 2: hd.nrows = std::atoi(v); break;
      case 3: hd.xll = std::atof(v); break;
      case 4: hd.yll = std::atof(v); break;
      case 5: hd.cellsize = std::atof(v); break;
      case 6: hd.nodata = std::atoi(v); break;
    }
  }
  return hd;
}

// src/dem_utils.h:237-295
DemHeader read_dem_header_binary(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) throw InputError("Error reading file: " + path);
  double a[6] = {0, 0, 0, 0, 0, 0};
  f.read(reinterpret_cast<char*>(a), sizeof(a));
  DemHeader hd;
  hd.ncols = int(a[0]); hd.nrows = int(a[1]);
  hd.xll = a[2]; hd.yll = a[3]; hd.cellsize = a[4]; hd.nodata = int(a[5]);
  return hd;
}

// src/matrix.h:1011-1038 (dims) + src/matrix_io.h:78-224 (values).
std::vector<double> read_ascii_raster(const std::string& path, int nrows, int ncols,
                                      int header_lines, bool check_dims) {
  std::string text = read_file(path);
  std::vector<double> out(size_t(nrows) * size_t(ncols), 0.0);
  size_t pos = 0;
  int line_no = 0, row = 0, first_tokens = -1, lines_seen = 0;
  while (pos < text.size()) {
    size_t nl = text.find('\n', pos);
    if (nl == std::string::npos) nl = text.size();
    std::string line = text.substr(pos, nl - pos);
    pos = nl + 1;
    line_no++;
    lines_seen++;
    if (line_no <= header_lines) continue;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    auto toks = split(line, ' ');
    if (first_tokens < 0) first_tokens = int(toks.size());
    if (toks.empty()) { row++; continue; }
    if (row >= nrows) throw InputError("Too many rows in " + path);
    if (int(toks.size()) > ncols)
      throw InputError("Error reading file: " + path + ". More than one space as separator?. Check row " +
                       std::to_string(row + 1));
    double* dst = out.data() + size_t(row) * ncols;
    for (size_t j = 0; j < toks.size(); j++) dst[j] = token_value(toks[j]);
    row++;
  }
  if (check_dims) {
    int nr_file = lines_seen - header_lines;
    if (nr_file != nrows || first_tokens != ncols)
      throw InputError("Invalid dimension of " + path + ". (row,col): (" + std::to_string(nr_file) +
                       "," + std::to_string(first_tokens) + ")");
  }
  return out;
}

// src/matrix_io.h:226-290
std::vector<double> read_binary_raster(const std::string& path, int nrows, int ncols,
                                       int header_values, bool check_header) {
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) throw InputError("Error reading file: " + path);
  std::vector<double> hdr(size_t(std::max(header_values, 0)));
  f.read(reinterpret_cast<char*>(hdr.data()), std::streamsize(hdr.size() * sizeof(double)));
  if (check_header && header_values >= 2 && (int(hdr[0]) != nrows || int(hdr[1]) != ncols))
    throw InputError("Invalid Matrix dimensions in " + path);
  std::vector<double> out(size_t(nrows) * size_t(ncols), 0.0);
  f.read(reinterpret_cast<char*>(out.data()), std::streamsize(out.size() * sizeof(double)));
  if (!f) throw InputError("Short read in " + path);
  return out;
}

void write_dem_ascii(const std::string& path, const DemHeader& hd, const double* z) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) throw InputError("Cannot write " + path);
  std::fprintf(fp, "ncols %d\nnrows %d\nxllcorner %.10f\nyllcorner %.10f\ncellsize %.10f\nNODATA_value %d\n",
               hd.ncols, hd.nrows, hd.xll, hd.yll, hd.cellsize, hd.nodata);
  for (int i = 0; i < hd.nrows; i++) {
    for (int j = 0; j < hd.ncols; j++)
      std::fprintf(fp, j ? " %.6f" : "%.6f", z[size_t(i) * hd.ncols + j]);
    std::fputc('\n', fp);
  }
  std::fclose(fp);
}

void write_raster_ascii(const std::string& path, int nrows, int ncols, const double* v,
                        int decimals) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) throw InputError("Cannot write " + path);
  for (int i = 0; i < nrows; i++) {
    for (int j = 0; j < ncols; j++)
      std::fprintf(fp, j ? " %.*f" : "%.*f", decimals, v[size_t(i) * ncols + j]);
    std::fputc('\n', fp);
  }
  std::fclose(fp);
}

void write_dem_binary(const std::string& path, const DemHeader& hd, const double* z) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw InputError("Cannot write " + path);
  double a[6] = {double(hd.ncols), double(hd.nrows), hd.xll, hd.yll, hd.cellsize, double(hd.nodata)};
  std::fwrite(a, sizeof(double), 6, fp);
  std::fwrite(z, sizeof(double), size_t(hd.nrows) * size_t(hd.ncols), fp);
  std::fclose(fp);
}

void write_raster_binary(const std::string& path, int nrows, int ncols, const double* v) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw InputError("Cannot write " + path);
  double a[2] = {double(nrows), double(ncols)};
  std::fwrite(a, sizeof(double), 2, fp);
  std::fwrite(v, sizeof(double), size_t(nrows) * size_t(ncols), fp);
  std::fclose(fp);
}

// src/output_ascii.h:52-75 -- interior only, fixed with 6 decimals.
void write_output_ascii(const std::string& path, int rows, int cols, const double* a) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) throw InputError("Cannot write " + path);
  for (int i = 1; i < rows - 1; i++) {
    for (int j = 1; j < cols - 1; j++)
      std::fprintf(fp, j < cols - 2 ? "%.6f " : "%.6f", a[size_t(i) * cols + j]);
    std::fputc('\n', fp);
  }
  std::fclose(fp);
}

// src/output_binary.h:65-80 -- (rows, cols) as T, then interior rows.
void write_output_binary(const std::string& path, int rows, int cols, const double* a) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw InputError("Cannot write " + path);
  double hdr[2] = {double(rows - 2), double(cols - 2)};
  std::fwrite(hdr, sizeof(double), 2, fp);
  for (int i = 1; i < rows - 1; i++) std::fwrite(a + size_t(i) * cols + 1, sizeof(double), size_t(cols - 2), fp);
  std::fclose(fp);
}

void pad_and_finalize(Case& c, const std::vector<double>& dem_in,
                      const std::vector<double>& n_in, const std::vector<double>& h_in,
                      const std::vector<double>& qx_in, const std::vector<double>& qy_in,
                      const std::vector<int32_t>& rmap_in, const Table* hyg,
                      const Table* roff) {
  const Config& cfg = c.cfg;
  const int nr = c.dem_hdr.nrows, nc = c.dem_hdr.ncols;
  if (nr <= 0 || nc <= 0) throw InputError("DEM has no cells");
  if (cfg.time_increment_fixed)
    throw InputError("time_increment_fixed=1 is not supported (the reference leaves sqrt(h) unset in that mode)");
  c.rows = nr + 2;
  c.cols = nc + 2;
  c.dx = c.dem_hdr.cellsize;
  const int rows = c.rows, cols = c.cols;
  const size_t ncell = size_t(rows) * size_t(cols);
  if (ncell > size_t(2147483647)) throw InputError("grid exceeds 2^31 cells");

  // bed: ghost 0, then walls or copies (src/triton_init.h:209-217)
  c.dem = pad(dem_in, nr, nc, 0.0);
  if (!cfg.open_boundaries) set_walls(c.dem, rows, cols);
  else copy_into_ghosts(c.dem, rows, cols);
  // Manning: ghost copies, then squared (src/triton_init.h:219-237)
  if (!n_in.empty()) c.nman2 = pad(n_in, nr, nc, 0.0);
  else c.nman2.assign(ncell, 0.0), c.nman2 = pad(std::vector<double>(size_t(nr) * nc, 0.0 + cfg.const_mann), nr, nc, 0.0);
  copy_into_ghosts(c.nman2, rows, cols);
  for (auto& v : c.nman2) v = v * v;
  // initial state (src/triton_init.h:239-284, 1225-1240)
  auto state = [&](const std::vector<double>& in) {
    if (in.empty()) return std::vector<double>(ncell, 0.0);
    auto p = pad(in, nr, nc, 0.0);
    if (cfg.open_boundaries) copy_into_ghosts(p, rows, cols);
    return p;
  };
  c.h = state(h_in);
  c.qx = state(qx_in);
  c.qy = state(qy_in);

  // point sources (src/triton_init.h:723-830)
  c.src_pos.clear();
  if (cfg.num_sources > 0) {
    if (!hyg) throw InputError("num_sources > 0 needs a hydrograph");
    if (int(cfg.src_x.size()) < cfg.num_sources) throw InputError("fewer source points than num_sources");
    for (int i = 0; i < cfg.num_sources; i++) {
      int col = calc_col(cfg.src_x[i], c.dem_hdr.xll, c.dem_hdr.cellsize);
      int row = calc_row(cfg.src_y[i], c.dem_hdr.yll, c.dem_hdr.cellsize, nr);
      if (col >= nc || row >= nr || col < 0 || row < 0)
        throw InputError("Source " + std::to_string(i + 1) + " is out of bounds");
      c.src_pos.push_back((row + 1) * cols + (col + 1));
    }
    c.hyg_nrows = int(hyg->rows.size());
    c.hyg_time.resize(size_t(c.hyg_nrows));
    c.hyg_val.resize(size_t(c.hyg_nrows) * size_t(cfg.num_sources));
    for (int k = 0; k < c.hyg_nrows; k++) {
      const auto& r = hyg->rows[size_t(k)];
      c.hyg_time[size_t(k)] = r.at(0) * 3600.0;
      for (int s = 0; s < cfg.num_sources; s++) {
        if (size_t(s + 1) >= r.size()) throw InputError("hydrograph row has too few columns");
        c.hyg_val[size_t(k) * cfg.num_sources + s] = r[size_t(s + 1)];
      }
    }
    if (c.hyg_nrows < 2) throw InputError("hydrograph needs at least two rows");
  }
  // observation cells (src/triton_init.h:834-909).  The reference only
  // prints an error for a point outside the DEM and then reads outside its
  // arrays; here such a point is an input error.
  c.obs_pos.clear();
  for (size_t i = 0; i < cfg.obs_x.size(); i++) {
    int col = calc_col(cfg.obs_x[i], c.dem_hdr.xll, c.dem_hdr.cellsize);
    int row = calc_row(cfg.obs_y[i], c.dem_hdr.yll, c.dem_hdr.cellsize, nr);
    if (cfg.time_series_flag && (col < 0 || col >= nc || row < 0 || row >= nr))
      throw InputError("observation " + std::to_string(i + 1) + " is out of bounds");
    c.obs_pos.push_back((row + 1) * cols + (col + 1));
  }

  // external boundaries (src/triton_init.h:913-1055, src/extbc.h:110-230)
  c.bc_index.clear(); c.bc_type.clear(); c.bc_start.clear(); c.bc_nrows.clear();
  c.bc_var1.clear(); c.bc_var2.clear();
  for (int i = 0; i < cfg.num_extbc; i++) {
    int type = cfg.extbc_type[size_t(i)];
    std::vector<std::vector<double>> rowsv;
    if (type == 1) {
      Table t = load_table_csv(cfg.extbc_fname[size_t(i)]);
      for (auto& r : t.rows) {
        if (r.size() < 2) throw InputError("extbc level file row needs two columns");
        rowsv.push_back({r[0] * 3600.0, r[1]});
      }
    } else {
      double v = to_double(cfg.extbc_fname[size_t(i)]);
      rowsv.push_back({v, v});
    }
    int ec0 = calc_col(cfg.extbc_x1[size_t(i)], c.dem_hdr.xll, c.dem_hdr.cellsize);
    int er0 = calc_row(cfg.extbc_y1[size_t(i)], c.dem_hdr.yll, c.dem_hdr.cellsize, nr);
    int ec1 = calc_col(cfg.extbc_x2[size_t(i)], c.dem_hdr.xll, c.dem_hdr.cellsize);
    int er1 = calc_row(cfg.extbc_y2[size_t(i)], c.dem_hdr.yll, c.dem_hdr.cellsize, nr);
    if (ec0 >= nc || ec1 >= nc) throw InputError("extbc x-coordinates must be at the domain boundary");
    if (er0 >= nr || er1 >= nr) throw InputError("extbc y-coordinates must be at the domain boundary");
    if (ec0 != 0 && ec0 != nc - 1 && er0 != 0 && er0 != nr - 1) throw InputError("extbc point 1 not on the boundary");
    if (ec1 != 0 && ec1 != nc - 1 && er1 != 0 && er1 != nr - 1) throw InputError("extbc point 2 not on the boundary");
    if (ec0 != ec1 && er0 != er1) throw InputError("extbc points must be aligned with a boundary");
    int ncells = (ec0 == ec1) ? std::abs(er0 - er1) + 1 : std::abs(ec0 - ec1) + 1;
    int cmin = std::min(ec0, ec1), cmax = std::max(ec0, ec1);
    int rmin = std::min(er0, er1), rmax = std::max(er0, er1);
    int loc = -1;
    if (cmin == cmax) loc = (cmin == 0) ? 0 : 2;
    if (rmin == rmax) loc = (rmin == 0) ? 1 : 3;
    int start = int(c.bc_var1.size());
    for (int k = 0; k < ncells; k++) {
      int ic = -1, ir = -1;
      if (loc == 0) { ic = 0; ir = rmin + k; }
      if (loc == 2) { ic = nc - 1; ir = rmin + k; }
      if (loc == 1) { ic = cmin + k; ir = 0; }
      if (loc == 3) { ic = cmin + k; ir = nr - 1; }
      c.bc_index.push_back((ir + 1) * cols + ic + 1);
      c.bc_type.push_back(type);
      c.bc_start.push_back(start);
      c.bc_nrows.push_back(int(rowsv.size()));
    }
    for (auto& r : rowsv) { c.bc_var1.push_back(r[0]); c.bc_var2.push_back(r[1]); }
  }
  // ghosts beside external-boundary cells take the inner bed and state
  // (src/triton_init.h:1243-1285)
  for (int ii : c.bc_index) {
    int ix = ii / cols, iy = ii % cols;
    auto cp = [&](int g) { c.dem[size_t(g)] = c.dem[size_t(ii)]; c.h[size_t(g)] = c.h[size_t(ii)];
                           c.qx[size_t(g)] = c.qx[size_t(ii)]; c.qy[size_t(g)] = c.qy[size_t(ii)]; };
    if (iy == 1) cp(ii - 1);
    if (iy == cols - 2) cp(ii + 1);
    if (ix == 1) cp(ii - cols);
    if (ix == rows - 2) cp(ii + cols);
  }

  // runoff (src/triton_init.h:1299-1357 with src/inflow.h:257-297 units)
  c.num_runoffs = cfg.num_runoffs;
  c.runoff_id.assign(ncell, -1);
  c.roff_time.clear(); c.roff_rate.clear(); c.roff_nrows = 0; c.roff_row0 = 0;
  if (cfg.num_runoffs > 0) {
    if (!roff) throw InputError("num_runoffs > 0 needs a runoff file");
    c.roff_nrows = int(roff->rows.size());
    if (c.roff_nrows < 2) throw InputError("runoff file needs at least two rows");
    int nin = roff->ncols_value();
    std::vector<std::vector<double>> d = roff->rows;
    for (auto& r : d) r.at(0) = r[0] * 3600.0;
    for (int i = 0; i < nin; i++)
      for (int k = 0; k < c.roff_nrows; k++) {
        auto& r = d[size_t(k)];
        if (size_t(i + 1) < r.size()) r[size_t(i + 1)] /= 3600.0;
      }
    for (int i = 0; i < nin; i++)
      for (int k = 0; k < c.roff_nrows; k++) {
        auto& r = d[size_t(k)];
        if (size_t(i + 1) < r.size()) r[size_t(i + 1)] *= 0.001;
      }
    c.roff_time.resize(size_t(c.roff_nrows));
    for (int k = 0; k < c.roff_nrows; k++) c.roff_time[size_t(k)] = d[size_t(k)][0];
    c.roff_rate.assign(size_t(cfg.num_runoffs) * size_t(c.roff_nrows), 0.0);
    for (int z = 0; z < cfg.num_runoffs; z++)
      for (int k = 0; k < c.roff_nrows; k++) {
        const auto& r = d[size_t(k)];
        if (size_t(z + 1) >= r.size()) throw InputError("runoff row has fewer zones than num_runoffs");
        c.roff_rate[size_t(z) * c.roff_nrows + k] = r[size_t(z + 1)];
      }
    double t0 = cfg.sim_start_time;
    for (int j = 0; j < c.roff_nrows - 1; j++)
      if (t0 > c.roff_time[size_t(j)] && t0 <= c.roff_time[size_t(j + 1)]) c.roff_row0 = j;
    if (rmap_in.empty()) throw InputError("num_runoffs > 0 needs a runoff_map");
    std::fill(c.runoff_id.begin(), c.runoff_id.end(), 0);
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < nc; j++) {
        int32_t z = rmap_in[size_t(i) * nc + j];
        if (z < 0 || z >= cfg.num_runoffs) throw InputError("runoff_map zone id out of range");
        c.runoff_id[size_t(i + 1) * cols + j + 1] = z;
      }
  }

  // maximum depth starts as the initial depth (src/triton_init.h:1241)
  c.maxh = c.h;

  // all-dry fallback dt (src/triton_simulate.h:163-187)
  double idt = cfg.time_series_flag ? std::fmin(cfg.print_observation, cfg.print_interval)
                                    : cfg.print_interval;
  if (cfg.num_runoffs > 0) idt = std::fmin(idt, c.roff_time[1] - c.roff_time[0]);
  if (cfg.num_sources > 0) idt = std::fmin(idt, c.hyg_time[1] - c.hyg_time[0]);
  c.init_dt = idt * 0.01;
}

Case load_case(const std::string& cfg_path, const std::string& base_dir) {
  Case c;
  c.cfg = load_config(cfg_path, base_dir);
  const Config& cfg = c.cfg;
  // input_option=PAR (src/triton_init.h:397-563) reads rasters split per
  // rank as <name>_<rank>.<ext>, the DEM header from header_filename (ASCII
  // whatever the input format) and the DEM itself without it.  One GPU holds
  // the whole grid, so it reads the one-partition set, rank 00.
  const bool par = cfg.input_option == "PAR";
  if (!par && cfg.input_option != "SEQ") throw InputError("input_option must be SEQ or PAR");
  if (cfg.dem_filename.empty()) throw InputError("dem_filename is required");
  if (par && cfg.header_filename.empty()) throw InputError("input_option=PAR needs header_filename");
  auto part = [&](const std::string& name, const char* ext) { return par ? name + "_00" + ext : name; };
  const bool asc = cfg.input_format == "ASC";
  if (par) c.dem_hdr = read_dem_header_ascii(cfg.header_filename);
  else c.dem_hdr = asc ? read_dem_header_ascii(cfg.dem_filename) : read_dem_header_binary(cfg.dem_filename);
  const int nr = c.dem_hdr.nrows, nc = c.dem_hdr.ncols;
  if (nr <= 0 || nc <= 0) throw InputError("bad DEM header in " + (par ? cfg.header_filename : cfg.dem_filename));
  const std::string dem_path = part(cfg.dem_filename, ".dem");
  std::vector<double> dem;
  if (par) dem = asc ? read_ascii_raster(dem_path, nr, nc, 0, false) : read_binary_raster(dem_path, nr, nc, 2, true);
  else dem = asc ? read_ascii_raster(dem_path, nr, nc, 6, false) : read_binary_raster(dem_path, nr, nc, 6, false);
  std::vector<double> nin, hin, qxin, qyin, mhin;
  std::vector<int32_t> rmap;
  if (!cfg.n_infile.empty()) nin = load_raster(cfg, part(cfg.n_infile, ".mann"), nr, nc);
  if (cfg.checkpoint_id > 0) {
    // hotstart from the BIN rasters of a previous print (src/triton_init.h:303-353,
    // :576-631), the running maximum too when it is printed
    char tag[32];
    std::snprintf(tag, sizeof(tag), "%02d_00.out", cfg.checkpoint_id);
    std::string bin = resolve(cfg.output_folder, base_dir) + "/bin/";
    hin = read_binary_raster(bin + "H_" + tag, nr, nc, 2, true);
    qxin = read_binary_raster(bin + "QX_" + tag, nr, nc, 2, true);
    qyin = read_binary_raster(bin + "QY_" + tag, nr, nc, 2, true);
    if (!cfg.max_value_print_option.empty()) mhin = read_binary_raster(bin + "MH_" + tag, nr, nc, 2, true);
  } else {
    if (!cfg.h_infile.empty()) hin = load_raster(cfg, part(cfg.h_infile, ".out"), nr, nc);
    if (!cfg.qx_infile.empty()) qxin = load_raster(cfg, part(cfg.qx_infile, ".out"), nr, nc);
    if (!cfg.qy_infile.empty()) qyin = load_raster(cfg, part(cfg.qy_infile, ".out"), nr, nc);
  }
  if (!cfg.runoff_map.empty() && cfg.num_runoffs > 0) {
    const std::string rmap_path = part(cfg.runoff_map, ".rmap");
    if (asc) {
      auto v = read_ascii_raster(rmap_path, nr, nc, 0, true);
      rmap.resize(v.size());
      for (size_t i = 0; i < v.size(); i++) rmap[i] = int32_t(v[i]);
    } else {
      // matrix<int> BIN: int32 header (rows, cols) then int32 data
      std::ifstream f(rmap_path, std::ios::binary);
      if (!f.good()) throw InputError("Error reading file: " + rmap_path);
      int32_t hdr[2] = {0, 0};
      f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
      if (hdr[0] != nr || hdr[1] != nc) throw InputError("Invalid Matrix dimensions in " + rmap_path);
      rmap.resize(size_t(nr) * nc);
      f.read(reinterpret_cast<char*>(rmap.data()), std::streamsize(rmap.size() * sizeof(int32_t)));
      if (!f) throw InputError("Short read in " + rmap_path);
    }
  }
  Table hyg, roff;
  if (cfg.num_sources > 0) hyg = load_table_csv(cfg.hydrograph_filename);
  if (cfg.num_runoffs > 0) roff = load_table_csv(cfg.runoff_filename);
  pad_and_finalize(c, dem, nin, hin, qxin, qyin, rmap, cfg.num_sources > 0 ? &hyg : nullptr,
                   cfg.num_runoffs > 0 ? &roff : nullptr);
  if (!mhin.empty()) {
    c.maxh = pad(mhin, nr, nc, 0.0);
    if (cfg.open_boundaries) copy_into_ghosts(c.maxh, c.rows, c.cols);
  }
  return c;
}

}  // namespace tr
