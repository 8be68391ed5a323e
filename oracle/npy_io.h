// Minimal .npy reader/writer and the node-keyed input generator shared by the
// oracle drivers.  Test infrastructure only: nothing in the product links this.
//
// The input generator defines "x_i ~ U[-1,1] per CG node from seed s" keyed by
// the node's integer coordinates (not by its position in a vector), so the same
// input can be produced in any node order: here, in the Python tests (numpy
// restatement in tests/golden/keyed_rng.py) and in the product's host code.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <stdexcept>

namespace npyio {

inline std::string dtype_str(const uint32_t *) { return "<u4"; }
inline std::string dtype_str(const int32_t *) { return "<i4"; }
inline std::string dtype_str(const double *) { return "<f8"; }
inline std::string dtype_str(const uint64_t *) { return "<u8"; }

template <typename T>
void save(const std::string &path, const T *data, const std::vector<size_t> &shape) {
  std::string hdr = "{'descr': '" + dtype_str(data) + "', 'fortran_order': False, 'shape': (";
  size_t n = 1;
  for (size_t i = 0; i < shape.size(); i++) {
    hdr += std::to_string(shape[i]);
    hdr += (shape.size() == 1 || i + 1 < shape.size()) ? "," : "";
    if (i + 1 < shape.size()) hdr += " ";
    n *= shape[i];
  }
  hdr += "), }";
  size_t total = 10 + hdr.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  hdr += std::string(pad, ' ');
  hdr += "\n";
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  const char magic[] = "\x93NUMPY\x01\x00";
  fwrite(magic, 1, 8, f);
  uint16_t hl = (uint16_t)hdr.size();
  fwrite(&hl, 2, 1, f);
  fwrite(hdr.data(), 1, hdr.size(), f);
  if (n) fwrite(data, sizeof(T), n, f);
  fclose(f);
}

// Reads a C-order little-endian array of the requested element type.
template <typename T>
std::vector<T> load(const std::string &path, std::vector<size_t> &shape) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot read " + path);
  char magic[8];
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "\x93NUMPY", 6) != 0)
    throw std::runtime_error("not an npy file: " + path);
  uint32_t hl = 0;
  if (magic[6] == 1) { uint16_t h16; if (fread(&h16, 2, 1, f) != 1) throw std::runtime_error("bad npy"); hl = h16; }
  else { if (fread(&hl, 4, 1, f) != 1) throw std::runtime_error("bad npy"); }
  std::string hdr(hl, ' ');
  if (fread(&hdr[0], 1, hl, f) != hl) throw std::runtime_error("bad npy header");
  T *probe = nullptr;
  if (hdr.find(dtype_str(probe)) == std::string::npos)
    throw std::runtime_error("unexpected dtype in " + path + ": " + hdr);
  if (hdr.find("'fortran_order': False") == std::string::npos)
    throw std::runtime_error("fortran order not supported: " + path);
  size_t p = hdr.find("'shape': (");
  p += 10;
  shape.clear();
  size_t n = 1;
  while (hdr[p] != ')') {
    while (hdr[p] == ' ' || hdr[p] == ',') p++;
    if (hdr[p] == ')') break;
    size_t v = 0;
    while (hdr[p] >= '0' && hdr[p] <= '9') v = v * 10 + (hdr[p++] - '0');
    shape.push_back(v);
    n *= v;
  }
  std::vector<T> out(n);
  if (n && fread(out.data(), sizeof(T), n, f) != n) throw std::runtime_error("short npy: " + path);
  fclose(f);
  return out;
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// U[-1,1) value for the node with integer coordinates c[0..dim-1].
inline double keyed_uniform(uint64_t seed, const uint32_t *c, int dim) {
  uint64_t s = mix64(seed + 0x9E3779B97F4A7C15ull);
  for (int d = 0; d < dim; d++)
    s = mix64(s ^ ((uint64_t)c[d] * 0x9E3779B97F4A7C15ull + (uint64_t)d));
  return 2.0 * ((double)(s >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

}  // namespace npyio
