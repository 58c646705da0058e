// oracle_count.cpp -- operation-counting build of the oracle (test and
// measurement infrastructure only).
//
// Compiles ocfd_oracle.c unchanged with `double` replaced by a wrapper type
// whose arithmetic operators count every floating-point operation the
// restatement executes.  Because the oracle follows the reference's
// expressions one for one, the counts are the reference algorithm's own
// operation counts (constant sub-expressions folded at compile time, such as
// 13.d0/12.d0, are not counted, as in the compiled reference).
//
//   oracle_count_reset()      zero the counters
//   oracle_count_read(out[5]) add/sub, mul, div, sqrt, pow executed since reset
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ocfd_hip.h"

namespace {
int64_t g_cnt[5];
}

struct cd {
  double v;
  cd() = default;
  cd(double x) : v(x) {}
  explicit operator double() const { return v; }
};
static inline cd operator+(cd a, cd b) { g_cnt[0]++; return cd(a.v + b.v); }
static inline cd operator-(cd a, cd b) { g_cnt[0]++; return cd(a.v - b.v); }
static inline cd operator*(cd a, cd b) { g_cnt[1]++; return cd(a.v * b.v); }
static inline cd operator/(cd a, cd b) { g_cnt[2]++; return cd(a.v / b.v); }
static inline cd operator-(cd a) { return cd(-a.v); }
static inline bool operator<(cd a, cd b) { return a.v < b.v; }
static inline bool operator>(cd a, cd b) { return a.v > b.v; }
static inline bool operator<=(cd a, cd b) { return a.v <= b.v; }
static inline bool operator>=(cd a, cd b) { return a.v >= b.v; }
static inline bool operator==(cd a, cd b) { return a.v == b.v; }
static inline bool operator!=(cd a, cd b) { return a.v != b.v; }
static inline cd sqrt(cd a) { g_cnt[3]++; return cd(::sqrt(a.v)); }
static inline cd fabs(cd a) { return cd(::fabs(a.v)); }
static inline cd copysign(double a, cd b) { return cd(::copysign(a, b.v)); }
static inline cd pow(cd a, cd b) { g_cnt[4]++; return cd(::pow(a.v, b.v)); }

#define OCFD_ORACLE_H
typedef struct oracle_block oracle_block;
#define double cd
extern "C" {
#include "ocfd_oracle.c"
}
#undef double

extern "C" int oracle_count_reset(void) {
  memset(g_cnt, 0, sizeof g_cnt);
  return 0;
}
extern "C" int oracle_count_read(int64_t out[5]) {
  memcpy(out, g_cnt, sizeof g_cnt);
  return 0;
}
