// Text output of ion frames in the reference's formats (opendust.py:1046-1115):
// XYZ frames, per-frame CSV files and the RESTART file.  The reference formats
// every number with Python's "{}".format (csv.writer does the same), i.e. the
// shortest decimal string that reads back to the same double, in fixed
// notation for decimal exponents -4 <= e < 16 and scientific otherwise.  The
// shortest digits come from std::to_chars; the layout rules are Python's
// float repr.  Frames are formatted by several threads in ion chunks and
// written in order.

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "opendust_amd.h"

extern int od_set_error_message(int code, const char* msg);

namespace {

// Python repr of a double; returns the number of characters written (<= 32).
int repr_double(double v, char* out) {
  char* o = out;
  if (std::isnan(v)) {
    std::memcpy(o, "nan", 3);
    return 3;
  }
  if (std::signbit(v)) *o++ = '-';
  if (std::isinf(v)) {
    std::memcpy(o, "inf", 3);
    return (int)(o - out) + 3;
  }
  char buf[40];
  const auto r = std::to_chars(buf, buf + sizeof(buf), std::fabs(v), std::chars_format::scientific);
  // buf: d[.ddd]e[+-]xx
  char digits[24];
  int k = 0;
  const char* p = buf;
  digits[k++] = *p++;
  if (*p == '.') {
    ++p;
    while (*p != 'e') digits[k++] = *p++;
  }
  ++p;  // 'e'
  int e = 0;
  std::from_chars(p + (*p == '+' ? 1 : 0), r.ptr, e);
  const int decpt = e + 1;  // value = 0.d1d2..dk x 10^decpt
  if (decpt <= -4 || decpt > 16) {
    *o++ = digits[0];
    if (k > 1) {
      *o++ = '.';
      std::memcpy(o, digits + 1, k - 1);
      o += k - 1;
    }
    *o++ = 'e';
    int x = decpt - 1;
    *o++ = x < 0 ? '-' : '+';
    if (x < 0) x = -x;
    if (x < 10) *o++ = '0';
    const auto rr = std::to_chars(o, o + 4, x);
    o = rr.ptr;
  } else if (decpt <= 0) {
    *o++ = '0';
    *o++ = '.';
    for (int i = 0; i < -decpt; ++i) *o++ = '0';
    std::memcpy(o, digits, k);
    o += k;
  } else if (decpt >= k) {
    std::memcpy(o, digits, k);
    o += k;
    for (int i = k; i < decpt; ++i) *o++ = '0';
    *o++ = '.';
    *o++ = '0';
  } else {
    std::memcpy(o, digits, decpt);
    o += decpt;
    *o++ = '.';
    std::memcpy(o, digits + decpt, k - decpt);
    o += k - decpt;
  }
  return (int)(o - out);
}

void append_repr(std::string& s, double v) {
  char b[40];
  s.append(b, repr_double(v, b));
}

// Formats rows [0, n) with `row(i, string&)` on up to 16 threads, in order.
template <class Row>
std::vector<std::string> format_rows(int64_t n, Row&& row) {
  const int64_t per = 16384;
  const int64_t n_chunks = (n + per - 1) / per;
  std::vector<std::string> parts((size_t)std::max<int64_t>(n_chunks, 0));
  const int n_thr = (int)std::max<int64_t>(1, std::min<int64_t>(
      {n_chunks, (int64_t)std::max(1u, std::thread::hardware_concurrency()), (int64_t)16}));
  auto work = [&](int t) {
    for (int64_t c = t; c < n_chunks; c += n_thr) {
      std::string& s = parts[(size_t)c];
      s.reserve((size_t)per * 64);
      for (int64_t i = c * per; i < std::min(n, (c + 1) * per); ++i) row(i, s);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < n_thr; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  return parts;
}

int write_parts(const char* path, const char* mode, const std::string& head, const std::vector<std::string>& parts) {
  FILE* f = std::fopen(path, mode);
  if (!f) {
    std::string m = std::string("cannot open ") + path;
    return od_set_error_message(OD_ERR_ARG, m.c_str());
  }
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  for (const auto& s : parts) ok = ok && std::fwrite(s.data(), 1, s.size(), f) == s.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    std::string m = std::string("write to ") + path + " failed";
    return od_set_error_message(OD_ERR_ARG, m.c_str());
  }
  return OD_OK;
}

}  // namespace

extern "C" int od_write_xyz(const char* path, const double* pos, int64_t n, int64_t frame) {
  if (!path || (!pos && n > 0) || n < 0) return od_set_error_message(OD_ERR_ARG, "od_write_xyz: bad argument");
  std::string head = std::to_string(n) + "\n" + std::to_string(frame) + "\n";
  auto parts = format_rows(n, [&](int64_t i, std::string& s) {
    s += "Ar\t";
    append_repr(s, pos[3 * i]);
    s += '\t';
    append_repr(s, pos[3 * i + 1]);
    s += '\t';
    append_repr(s, pos[3 * i + 2]);
    s += '\n';
  });
  return write_parts(path, "ab", head, parts);
}

extern "C" int od_write_csv(const char* path, const double* pos, int64_t n, double density, double mass) {
  if (!path || (!pos && n > 0) || n < 0) return od_set_error_message(OD_ERR_ARG, "od_write_csv: bad argument");
  std::string tail;
  tail += ',';
  append_repr(tail, density);
  tail += ',';
  append_repr(tail, mass);
  tail += "\r\n";
  auto parts = format_rows(n, [&](int64_t i, std::string& s) {
    append_repr(s, pos[3 * i]);
    s += ',';
    append_repr(s, pos[3 * i + 1]);
    s += ',';
    append_repr(s, pos[3 * i + 2]);
    s += tail;
  });
  return write_parts(path, "wb", "x,y,z,density,mass\r\n", parts);
}

extern "C" int od_write_restart(const char* path, const double* pos, const double* vel, int64_t n) {
  if (!path || ((!pos || !vel) && n > 0) || n < 0)
    return od_set_error_message(OD_ERR_ARG, "od_write_restart: bad argument");
  auto parts = format_rows(n, [&](int64_t i, std::string& s) {
    for (int k = 0; k < 3; ++k) {
      append_repr(s, pos[3 * i + k]);
      s += '\t';
    }
    for (int k = 0; k < 3; ++k) {
      append_repr(s, vel[3 * i + k]);
      s += (k < 2) ? '\t' : '\n';
    }
  });
  return write_parts(path, "wb", std::string(), parts);
}
