// wb_flowsnake.h -- flowsnake tile enumeration (flowsnake.cpp:46-136), as
// integer-only host/device functions.
//
// toFlowsnake(n) = baseFlow(iToFlowsnake(n)): iToFlowsnake permutes the 11
// base-7 digits of n (offset to be non-negative) with a 6-state digit
// automaton; baseFlow reads the result as centred base-7 digits of an
// Eisenstein integer in base 2-w.  The inverse used to find the counter of a
// tile address runs the same two steps backwards: the least-norm remainder of
// x+yw modulo 2-w is the unique unit (or 0) congruent to x+2y (mod 7), since
// w == 2 (mod 2-w), and the automaton's rows are permutations of 0..6.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WB_FHD __host__ __device__ inline
#else
#define WB_FHD inline
#endif

namespace wb {

// forwardFlowsnakeTable, flowsnake.cpp:46-54: entry = (next state << 4) | digit
#define WB_FLOW_TABLE \
  {0x52, 0x05, 0x06, 0x24, 0x33, 0x40, 0x01}, {0x31, 0x10, 0x12, 0x05, 0x43, 0x54, 0x16}, \
  {0x46, 0x24, 0x21, 0x10, 0x53, 0x35, 0x22}, {0x31, 0x10, 0x03, 0x54, 0x36, 0x35, 0x22}, \
  {0x46, 0x24, 0x13, 0x35, 0x42, 0x40, 0x01}, {0x52, 0x05, 0x23, 0x40, 0x51, 0x54, 0x16}

constexpr int64_t FLOW_IN_OFFSET = 1235829214;  // flowsnake.cpp:96
constexpr int64_t FLOW_OUT_OFFSET = 988663371;  // flowsnake.cpp:109

WB_FHD int flow_digit_next(int state, int d)
{
  const unsigned char t[6][7] = {WB_FLOW_TABLE};
  return t[state][d];
}

// iToFlowsnake
WB_FHD int64_t i_to_flowsnake(int64_t n)
{
  int dig[11];
  int64_t v = n + FLOW_IN_OFFSET;
  for (int i = 0; i < 11; i++) { dig[i] = (int)(v % 7); v /= 7; }
  int st = 0;
  v = 0;
  for (int i = 10; i >= 0; i--) {
    int e = flow_digit_next(st, dig[i]);
    st = e >> 4;
    v = 7 * v + (e & 7);
  }
  return v - FLOW_OUT_OFFSET;
}

// inverse of iToFlowsnake
WB_FHD int64_t flowsnake_to_i(int64_t m)
{
  int dig[11];
  int64_t v = m + FLOW_OUT_OFFSET;
  if (v < 0) return INT64_MIN;
  for (int i = 0; i < 11; i++) { dig[i] = (int)(v % 7); v /= 7; }
  if (v != 0) return INT64_MIN;
  int st = 0;
  v = 0;
  for (int i = 10; i >= 0; i--) {
    int din = -1;
    for (int d = 0; d < 7; d++)
      if ((flow_digit_next(st, d) & 7) == dig[i]) { din = d; break; }
    int e = flow_digit_next(st, din);
    st = e >> 4;
    v = 7 * v + din;
  }
  return v - FLOW_IN_OFFSET;
}

// baseFlow: centred base-7 integer -> Eisenstein (x + y w) in base 2-w
WB_FHD void base_flow(int64_t m, int *ex, int *ey)
{
  int64_t rx = 0, ry = 0, px = 1, py = 0;
  while (m) {
    int64_t d = m % 7;
    if (d > 3) d -= 7;
    if (d < -3) d += 7;
    m = (m - d) / 7;
    int64_t dy = (d + 4) / 3 - 1, dx = d - 2 * dy;  // baseFlowDig
    // (rx + ry w) += (dx + dy w)(px + py w); w^2 = -1 - w
    rx += dx * px - dy * py;
    ry += dx * py + dy * px - dy * py;
    // (px + py w) *= (2 - w)
    int64_t nx = 2 * px + py, ny = -px + py * 3;
    px = nx;
    py = ny;
  }
  *ex = (int)rx;
  *ey = (int)ry;
}

// baseSeven (flowsnake.cpp:75-90)
WB_FHD int64_t base_seven(int64_t x, int64_t y)
{
  int64_t ret = 0, p7 = 1;
  while (x * x + y * y - x * y) {
    int64_t d = ((x + 2 * y) % 7 + 7) % 7;
    if (d > 3) d -= 7;
    int64_t dy = (d + 4) / 3 - 1, dx = d - 2 * dy;
    x -= dx;
    y -= dy;
    int64_t nx = (3 * x - y) / 7, ny = (x + 2 * y) / 7;  // (x + y w)/(2 - w)
    x = nx;
    y = ny;
    ret += d * p7;
    p7 *= 7;
  }
  return ret;
}

WB_FHD void to_flowsnake_dev(int64_t n, int *ex, int *ey) { base_flow(i_to_flowsnake(n), ex, ey); }

WB_FHD bool from_flowsnake_dev(int ex, int ey, int lo, int hi, int *n)
{
  int64_t m = base_seven(ex, ey);
  int64_t v = flowsnake_to_i(m);
  if (v == INT64_MIN || v < lo || v > hi) return false;
  *n = (int)v;
  return true;
}

}  // namespace wb
