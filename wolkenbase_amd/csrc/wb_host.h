// wb_host.h -- host-side set-up shared by libwb_gpu (and by the host build
// of the device math used in CPU tests).  Everything here runs on the CPU
// with the same glibc calls the reference makes, so its results are the
// reference's bit for bit.
#pragma once
#include <stdint.h>
#include "wb_math.h"

namespace wb {

// fillTanTables (angle.cpp:305-320): tan over [-45deg,45deg) in 511 steps,
// sin/cos at the half-step offsets, from x87 sinl/cosl/tanl.
void fill_atan_tables(double *tan511, double *cos512, double *sin512);

// K = 0x40000000/M_PIl of angle.cpp:151 as an exact double pair and as the
// x87 significand/exponent.
AtanConst atan_const();

// Octree::sizeFit (octree.cpp:268-310) over a list of corner points.
void size_fit(const double *pts, int npts, double center[3], double *side);

// Flowsnake tiling parameters: BoundRect of the header corners at
// orientation 0 (boundrect.cpp:61-74, dirbound point.cpp:64-72), the cube of
// wolkencanvas.cpp:511-518 and Flowsnake::setSize (flowsnake.cpp:208-229).
struct SnakeSetup {
  double cx, cy, spacing, radius;
  int lo, hi, index;
};
SnakeSetup snake_setup(const double bbmin[3], const double bbmax[3], double tile_size);
// the same over the corners of several LAS headers
SnakeSetup snake_setup_corners(const double *corners, int ncorners, double tile_size);
// BoundRect at orientation 0 of a list of corners (x, y, z each):
// left, bottom, right, top, low, high (boundrect.cpp:53-70, boundrect.h:41-64)
void bound_rect(const double *corners, int ncorners, double lbrthl[6]);

// toFlowsnake(n) (flowsnake.cpp:92-136) -> Eisenstein (x, y)
void to_flowsnake(int n, int *ex, int *ey);
// inverse: Eisenstein (x, y) -> counter n (returns false if outside
// [lo, hi] of the given setup)
bool from_flowsnake(int ex, int ey, int lo, int hi, int *n);

}  // namespace wb
