// Output files in the formats of the reference's solver_output
// (solver_output.cu:14-316).  Host code only; the arrays are the host views a
// download fills (original particle order).
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <vector>

#include "../../include/gdem.h"

extern "C" {

// solver_output_init (solver_output.cu:14-41): the folder exists afterwards.
int gdem_output_init(const char* dir) {
  if (!dir) return GDEM_E_ARG;
  struct stat st;
  if (stat(dir, &st) == 0) return S_ISDIR(st.st_mode) ? GDEM_OK : GDEM_E_ARG;
  if (mkdir(dir, 0755) == 0 || errno == EEXIST) return GDEM_OK;
  return GDEM_E_ARG;
}

// write_frame_bin (solver_output.cu:259-316): dir/result_%012d.bin holding
// [int32 active count][float32 xyz of each active particle][float32 radius
// of each], particles in id order, values scaled by length_factor before the
// conversion to float.
int gdem_write_frame_bin(const char* dir, int step, int n, const double* x, const double* r,
                         const int* is_active, double length_factor) {
  if (!dir || n < 0 || (n > 0 && (!x || !r || !is_active))) return GDEM_E_ARG;
  char name[4096];
  std::snprintf(name, sizeof(name), "%s/result_%012d.bin", dir, step);
  FILE* fp = std::fopen(name, "wb");
  if (!fp) return GDEM_E_ARG;
  int act = 0;
  for (int i = 0; i < n; ++i) act += is_active[i] == 1 ? 1 : 0;
  std::vector<float> pos((size_t)act * 3), rad((size_t)act);
  int k = 0;
  for (int i = 0; i < n; ++i) {
    if (is_active[i] == 0) continue;
    pos[(size_t)k * 3 + 0] = (float)(x[(size_t)i * 3 + 0] * length_factor);
    pos[(size_t)k * 3 + 1] = (float)(x[(size_t)i * 3 + 1] * length_factor);
    pos[(size_t)k * 3 + 2] = (float)(x[(size_t)i * 3 + 2] * length_factor);
    rad[k] = (float)(r[i] * length_factor);
    ++k;
  }
  bool ok = std::fwrite(&act, sizeof(int), 1, fp) == 1;
  if (act) {
    ok = ok && std::fwrite(pos.data(), sizeof(float), pos.size(), fp) == pos.size();
    ok = ok && std::fwrite(rad.data(), sizeof(float), rad.size(), fp) == rad.size();
  }
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? GDEM_OK : GDEM_E_ARG;
}

// The same file from the compacted arrays gdem_scene_frame produces on the
// device (n_active float32 xyz triples and radii, already scaled).
int gdem_write_frame_compact(const char* dir, int step, int n_active, const float* xyz, const float* r) {
  if (!dir || n_active < 0 || (n_active > 0 && (!xyz || !r))) return GDEM_E_ARG;
  char name[4096];
  std::snprintf(name, sizeof(name), "%s/result_%012d.bin", dir, step);
  FILE* fp = std::fopen(name, "wb");
  if (!fp) return GDEM_E_ARG;
  const size_t n = (size_t)n_active;
  bool ok = std::fwrite(&n_active, sizeof(int), 1, fp) == 1;
  if (n) {
    ok = ok && std::fwrite(xyz, sizeof(float), 3 * n, fp) == 3 * n;
    ok = ok && std::fwrite(r, sizeof(float), n, fp) == n;
  }
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? GDEM_OK : GDEM_E_ARG;
}

// write_header_text (solver_output.cu:62-85): dir/pid<pid>.txt created with
// its column header.
int gdem_write_pid_header(const char* dir, int pid) {
  if (!dir) return GDEM_E_ARG;
  char name[4096];
  std::snprintf(name, sizeof(name), "%s/pid%d.txt", dir, pid);
  FILE* fp = std::fopen(name, "w");
  if (!fp) return GDEM_E_ARG;
  std::fprintf(fp,
               "0_time(s),1_pid,2_x(m),3_y(m),4_z(m),5_vx(m/s),6_vy(m/s),7_vz(m/s),8_ax(m/s2),9_ay(m/s2),"
               "10_az(m/s2),11_angvx(rad/s),12_angvy(rad/s),13_angvz(rad/s),14_angax(rad/s2),15_angay(rad/s2),"
               "16_angaz(rad/s2),17_isActive\n");
  return std::fclose(fp) == 0 ? GDEM_OK : GDEM_E_ARG;
}

// write_single_text (solver_output.cu:112-172) with unit length and time
// factors: one row appended to dir/pid<pid>.txt for particle pid at `step`
// (time = dt * step).  a and anga are the particle's linear and angular
// accelerations of that step.
int gdem_write_pid_row(const char* dir, int step, double dt, int pid, const double x[3], const double v[3],
                       const double a[3], const double angv[3], const double anga[3], int is_active) {
  if (!dir || !x || !v || !a || !angv || !anga) return GDEM_E_ARG;
  char name[4096];
  std::snprintf(name, sizeof(name), "%s/pid%d.txt", dir, pid);
  FILE* fp = std::fopen(name, "a");
  if (!fp) return GDEM_E_ARG;
  const double time = dt * (double)step;
  std::fprintf(fp, "%f %d %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %d\n", time, pid, x[0], x[1], x[2], v[0],
               v[1], v[2], a[0], a[1], a[2], angv[0], angv[1], angv[2], anga[0], anga[1], anga[2], is_active);
  return std::fclose(fp) == 0 ? GDEM_OK : GDEM_E_ARG;
}

// write_initialPos_csv (solver_output.cu:87-110): rows appended to
// dir/initialPos.csv, "<id+1> 1 <diameter> 1000. <x> <y> <z>".
int gdem_write_initial_pos(const char* dir, int n, const double* x, const double* r) {
  if (!dir || n < 0 || (n > 0 && (!x || !r))) return GDEM_E_ARG;
  char name[4096];
  std::snprintf(name, sizeof(name), "%s/initialPos.csv", dir);
  FILE* fp = std::fopen(name, "a");
  if (!fp) return GDEM_E_ARG;
  for (int i = 0; i < n; ++i)
    std::fprintf(fp, "%d 1 %f 1000. %f %f %f\n", i + 1, r[i] * 2.0, x[(size_t)i * 3], x[(size_t)i * 3 + 1],
                 x[(size_t)i * 3 + 2]);
  return std::fclose(fp) == 0 ? GDEM_OK : GDEM_E_ARG;
}

}  // extern "C"
