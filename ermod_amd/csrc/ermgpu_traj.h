// ermgpu_traj.h -- host-side frame ingestion: DCD trajectories
// (trajectory-formats/dcd.f90:39-105), the PermIndex permutation and the
// SysWght frame weights (getconf_parallel / read_weight,
// setconf.F90:860-996).  Plain C++ on the host; the device side is the
// scatter kernel k_frame_scatter and erm_traj_* in ermgpu.hip.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace erm {

// A file of Fortran unformatted sequential records: each record is framed by
// a 4-byte length before and after it, in the byte order of the writer.
struct RecordFile {
  FILE *f = nullptr;
  bool swap = false;

  static uint32_t bswap(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
  }
  bool read_u32(uint32_t &v) {
    if (fread(&v, 4, 1, f) != 1) return false;
    if (swap) v = bswap(v);
    return true;
  }
  // the next record's payload; false at the end of the file or on a framing error
  bool record(std::vector<unsigned char> &buf) {
    uint32_t n = 0, m = 0;
    if (!read_u32(n)) return false;
    buf.resize(n);
    if (n && fread(buf.data(), 1, n, f) != n) return false;
    if (!read_u32(m) || m != n) return false;
    return true;
  }
  bool skip_record() {
    uint32_t n = 0, m = 0;
    if (!read_u32(n)) return false;
    if (fseek(f, (long)n, SEEK_CUR) != 0) return false;
    return read_u32(m) && m == n;
  }
};

// DCD: an 84-byte header record that starts with the 'CORD' magic (21
// int32, icntrl(11) = 1 when every frame carries a unit-cell record), a
// title record and an atom-count record; then per frame the optional cell
// record (6 doubles) and three records of FP32 x, y and z.
struct Dcd {
  RecordFile rf;
  bool have_cell = false;
  int32_t natom_file = 0;
  std::vector<unsigned char> buf;

  // 0 on success, else a message
  const char *open(const char *path) {
    rf.f = fopen(path, "rb");
    if (!rf.f) return "cannot open the DCD file";
    uint32_t n = 0;
    if (fread(&n, 4, 1, rf.f) != 1) return "empty DCD file";
    if (n != 84) {
      if (RecordFile::bswap(n) != 84) return "incorrect dcd format (header record)";
      rf.swap = true;
    }
    rewind(rf.f);
    if (!rf.record(buf) || buf.size() != 84) return "incorrect dcd format (header record)";
    int32_t h[21];
    memcpy(h, buf.data(), sizeof h);
    if (rf.swap)
      for (int i = 0; i < 21; ++i) h[i] = (int32_t)RecordFile::bswap((uint32_t)h[i]);
    // 'CORD' as a 4-byte integer in either byte order (dcd.f90:45-51)
    if (h[0] != 0x44524f43 && h[0] != 0x434f5244) return "incorrect dcd format";
    have_cell = (h[11] == 1);
    if (!rf.skip_record()) return "incorrect dcd format (title record)";
    if (!rf.record(buf) || buf.size() < 4) return "incorrect dcd format (atom-count record)";
    uint32_t na;
    memcpy(&na, buf.data(), 4);
    natom_file = (int32_t)(rf.swap ? RecordFile::bswap(na) : na);
    return nullptr;
  }
  void close() {
    if (rf.f) fclose(rf.f);
    rf.f = nullptr;
  }
  // read_trajectory (dcd.f90:67-105): 1 at the end of the file, -1 on a
  // malformed frame, 0 on success.  xyz: SoA x[natom], y[natom], z[natom];
  // the first natom values of each record are taken, as a Fortran read of
  // natom items from a longer record does.  cell: column-major 3x3 filled as
  // (1,1), (1,2), (2,2), (1,3), (2,3), (3,3) from the unit-cell record.
  int frame(int natom, bool periodic, float *xyz, double *cell) {
    for (int i = 0; i < 9; ++i) cell[i] = 0.0;
    if (periodic) {
      if (!rf.record(buf)) return feof(rf.f) ? 1 : -1;
      if (buf.size() < 48) return -1;
      double v[6];
      memcpy(v, buf.data(), 48);
      if (rf.swap)
        for (int i = 0; i < 6; ++i) {
          uint64_t u;
          memcpy(&u, &v[i], 8);
          u = ((uint64_t)RecordFile::bswap((uint32_t)u) << 32) | RecordFile::bswap((uint32_t)(u >> 32));
          memcpy(&v[i], &u, 8);
        }
      cell[0] = v[0];
      cell[3] = v[1];
      cell[4] = v[2];
      cell[6] = v[3];
      cell[7] = v[4];
      cell[8] = v[5];
    }
    for (int d = 0; d < 3; ++d) {
      if (!rf.record(buf)) return (d == 0 && !periodic && feof(rf.f)) ? 1 : -1;
      if (buf.size() < (size_t)natom * 4) return -1;
      memcpy(xyz + (size_t)d * natom, buf.data(), (size_t)natom * 4);
      if (rf.swap)
        for (int i = 0; i < natom; ++i) {
          uint32_t u;
          memcpy(&u, xyz + (size_t)d * natom + i, 4);
          u = RecordFile::bswap(u);
          memcpy(xyz + (size_t)d * natom + i, &u, 4);
        }
    }
    return 0;
  }
};

// PermIndex (setconf.F90:866-905): lines "i p" (1-based) setting the
// program atom p of trajectory atom i; unlisted atoms map to themselves.
// Every index must be in range and every target must appear exactly once.
// perm[i] is 0-based.  Returns 0, 1 when the file does not exist, -1 when it
// is ill-formed (set_pmt).
inline int read_permutation(const char *path, int n, int32_t *perm, std::string &msg) {
  FILE *f = fopen(path, "r");
  if (!f) return 1;
  for (int i = 0; i < n; ++i) perm[i] = i;
  long a, b;
  int rc = 0;
  char line[512];
  while (fgets(line, sizeof line, f)) {
    char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '\n' || *p == 0) continue;
    if (sscanf(p, "%ld %ld", &a, &b) != 2) break;  // list-directed read ends
    if (a < 1 || a > n || b < 1 || b > n) {
      msg = "PermIndex entry " + std::to_string(a) + " -> " + std::to_string(b) + " out of range (set_pmt)";
      rc = -1;
      break;
    }
    perm[a - 1] = (int32_t)(b - 1);
  }
  fclose(f);
  if (rc) return rc;
  std::vector<int> cnt(n, 0);
  for (int i = 0; i < n; ++i) ++cnt[perm[i]];
  for (int i = 0; i < n; ++i)
    if (cnt[i] != 1) {
      msg = "PermIndex: atom " + std::to_string(i + 1) + " has " + std::to_string(cnt[i]) + " entries (set_pmt)";
      return -1;
    }
  return 0;
}

// SysWght (read_weight, setconf.F90:955-986): one "index weight" line per
// frame read, starting over at the end of the file.
struct WeightFile {
  FILE *f = nullptr;
  bool open(const char *path) {
    f = fopen(path, "r");
    return f != nullptr;
  }
  void close() {
    if (f) fclose(f);
    f = nullptr;
  }
  bool read_one(double &w) {
    char line[512];
    while (fgets(line, sizeof line, f)) {
      char *p = line;
      while (*p == ' ' || *p == '\t') ++p;
      if (*p == '\n' || *p == 0) continue;
      long i;
      return sscanf(p, "%ld %lf", &i, &w) == 2;
    }
    return false;
  }
  // false when the file is ill-formed even from its start
  bool next(double &w) {
    if (read_one(w)) return true;
    rewind(f);
    return read_one(w);
  }
};

}  // namespace erm
