// case.hpp -- host-side model of one TRITON run as the reference's init leaves
// it: the .cfg keys, the ghost-padded SoA rasters and the forcing tables.
//
// Input formats follow the reference exactly:
//   .cfg key=value           src/config_utils.h:195-238 (parse), :422-517 (keys)
//   ESRI-ASCII DEM header    src/dem_utils.h:164-233; BIN DEM header :237-295
//   headerless ASC rasters   src/matrix_io.h:78-149 (atof only if '.' in token)
//   BIN rasters              src/matrix_io.h:226-260 (2-value T header)
//   .hyg / .roff CSV         src/inflow.h:176-219 (% comments), units :257-297
//   .src / .obs CSV          src/config_utils.h:241-302
//   .extbc CSV               src/config_utils.h:305-419, src/extbc.h:110-190
// Ghost ring, walls, n^2, source/runoff/extbc indices:
//   src/triton_init.h:195-355, 723-1055, 1059-1295, 1299-1357, 1595-1605.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace tr {

struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Parsed .cfg (the subset of arguments<T> the timestep and its I/O use).
struct Config {
  std::map<std::string, std::string> raw;
  std::string dem_filename, input_format, input_option, output_format, output_option;
  std::string output_folder, outfile_pattern, print_option, max_value_print_option;
  std::string hydrograph_filename, src_loc_file, runoff_filename, runoff_map;
  std::string observation_loc_file, extbc_file, extbc_dir, header_filename;
  std::string h_infile, qx_infile, qy_infile, n_infile;
  double const_mann = 0, time_step = 0, courant = 0, hextra = 0;
  double sim_start_time = 0, sim_duration = 0, print_interval = 0, print_observation = 0;
  int time_increment_fixed = 0, time_series_flag = 0, open_boundaries = 0;
  int num_sources = 0, num_runoffs = 0, num_extbc = 0;
  int checkpoint_id = 0, it_count = 0, it_print = 10000;
  std::vector<double> src_x, src_y, obs_x, obs_y;
  std::vector<int> extbc_type;
  std::vector<double> extbc_x1, extbc_y1, extbc_x2, extbc_y2;
  std::vector<std::string> extbc_fname;
};

// key=value parse with the reference's comment and quote rules.
std::map<std::string, std::string> parse_cfg_text(const std::string& text);
Config config_from_text(const std::string& text, const std::string& base_dir);
Config load_config(const std::string& cfg_path, const std::string& base_dir);

// Time-major table: rows of [t, v1, v2, ...] (hydrograph / runoff / extbc).
struct Table {
  std::vector<std::vector<double>> rows;
  int ncols_value() const { return rows.empty() ? 0 : int(rows.back().size()) - 1; }
};
Table load_table_csv(const std::string& path);

struct DemHeader {
  int ncols = 0, nrows = 0, nodata = 0;
  double xll = 0, yll = 0, cellsize = 0;
};
DemHeader read_dem_header_ascii(const std::string& path);
DemHeader read_dem_header_binary(const std::string& path);

// Raster readers return row-major nrows x ncols (no ghosts).
std::vector<double> read_ascii_raster(const std::string& path, int nrows, int ncols,
                                      int header_lines, bool check_dims);
std::vector<double> read_binary_raster(const std::string& path, int nrows, int ncols,
                                       int header_values, bool check_header);

// Writers used by the synthetic-input generators (reference formats).
void write_dem_ascii(const std::string& path, const DemHeader& hd, const double* z);
void write_raster_ascii(const std::string& path, int nrows, int ncols, const double* v,
                        int decimals);
void write_dem_binary(const std::string& path, const DemHeader& hd, const double* z);
void write_raster_binary(const std::string& path, int nrows, int ncols, const double* v);

// One model instance, padded (rows = nrows+2, cols = ncols+2).
struct Case {
  Config cfg;
  DemHeader dem_hdr;
  int rows = 0, cols = 0;
  double dx = 0;
  std::vector<double> h, qx, qy, dem, nman2, maxh;
  std::vector<int32_t> runoff_id;            // per cell (-1 without runoff)
  // forcing, converted to SI as the reference does at load
  std::vector<double> roff_time, roff_rate;  // rate zone-major [zone*nrows + k]
  int roff_nrows = 0, num_runoffs = 0, roff_row0 = 0;
  std::vector<double> hyg_time, hyg_val;     // val time-major [k*nsrc + s]
  int hyg_nrows = 0;
  std::vector<int32_t> src_pos;              // flat padded index per source
  std::vector<int32_t> obs_pos;
  std::vector<int32_t> bc_index, bc_type, bc_start, bc_nrows;
  std::vector<double> bc_var1, bc_var2;
  double init_dt = 0;                        // src/triton_simulate.h:163-187
};

// Build a Case from a .cfg (paths resolved against base_dir when relative).
Case load_case(const std::string& cfg_path, const std::string& base_dir);
// Finish a Case whose cfg and unpadded inputs were set in memory (generators).
void pad_and_finalize(Case& c, const std::vector<double>& dem_in,
                      const std::vector<double>& n_in, const std::vector<double>& h_in,
                      const std::vector<double>& qx_in, const std::vector<double>& qy_in,
                      const std::vector<int32_t>& rmap_in, const Table* hyg,
                      const Table* roff);

// Output writers (src/output_ascii.h:55-75, src/output_binary.h).
void write_output_ascii(const std::string& path, int rows, int cols, const double* padded);
void write_output_binary(const std::string& path, int rows, int cols, const double* padded);

}  // namespace tr
