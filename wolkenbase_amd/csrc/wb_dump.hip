// wb_dump.hip -- the octree as text: OctStore::dump (octree.cpp:888-891),
// Octree::dump (octree.cpp:360-376) and OctBuffer::dump (octree.cpp:673-689),
// with the numbers written by ldecimal (ldecimal.cpp:32-132).  One line per
// leaf in DFS order:
//   (cx,cy,cz)±half N points[, K stray points]
// then "T total points".  The per-leaf counts come from one kernel; the
// text is made on the host.
#include "../../include/wb_gpu.h"
#include "wb_internal.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

// per leaf: points with a finite location, and points inside the leaf cube
// (Cube::in, shape.cpp:38-43: every axis within side/2 of the centre)
__global__ void k_leaf_dump_counts(const uint32_t *leaf_start, const double *leaf_cube, uint32_t nl,
                                   const double *px, const double *py, const double *pz,
                                   uint32_t *nfinite, uint32_t *nin)
{
  const uint32_t l = blockIdx.x;
  if (l >= nl) return;
  const uint32_t b = leaf_start[l], e = leaf_start[l + 1];
  const double cx = leaf_cube[4 * (uint64_t)l], cy = leaf_cube[4 * (uint64_t)l + 1];
  const double cz = leaf_cube[4 * (uint64_t)l + 2], h = leaf_cube[4 * (uint64_t)l + 3] / 2;
  uint32_t f = 0, in = 0;
  for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    const double x = px[i], y = py[i], z = pz[i];
    f += (isfinite(x) && isfinite(y) && isfinite(z)) ? 1u : 0u;
    in += (fabs(x - cx) <= h && fabs(y - cy) <= h && fabs(z - cz) <= h) ? 1u : 0u;
  }
  for (int s = 32; s > 0; s >>= 1) {
    f += __shfl_xor(f, s, 64);
    in += __shfl_xor(in, s, 64);
  }
  if (threadIdx.x == 0) {
    nfinite[l] = f;
    nin[l] = in;
  }
}

// ldecimal (ldecimal.cpp:32-132): the shortest %e rendering that reads back
// within toler (exactly, for toler 0), then rewritten without an exponent
// where one is not needed
static std::string ldecimal(double x, double toler)
{
  char buffer[32], fmt[16];
  int iexp;
  if (toler > 0 && x != 0) {
    iexp = (int)floor(log10(fabs(x / toler)) - 1);
    if (iexp < 0) iexp = 0;
  } else {
    iexp = DBL_DIG - 1;
  }
  if (iexp > DBL_DIG) iexp = DBL_DIG + 1;
  // walk the digit count down while the text still reads back exactly, then
  // up until it does
  int step = -1, i = iexp;
  for (;;) {
    snprintf(fmt, sizeof fmt, "%%.%de", i);
    snprintf(buffer, sizeof buffer, fmt, x);
    const double back = atof(buffer);
    if (step > 0 && (fabs(x - back) <= toler || i >= DBL_DIG + 3)) break;
    if (fabs(x - back) > toler || i <= 0) step = 1;
    i += step;
  }
  char *dot = strchr(buffer, '.'), *ep = strchr(buffer, 'e');
  if (ep && !dot) {  // "2e+00" becomes "2.e+00"
    memmove(ep + 1, ep, buffer + 31 - ep);
    dot = ep++;
    *dot = '.';
  }
  if (!(dot && ep)) return buffer;  // inf, nan
  std::string m(buffer, dot - buffer), frac(dot + 1, ep - dot - 1), sign;
  if (m.size() > 1) {
    sign = m.substr(0, 1);
    m.erase(0, 1);
  }
  int e = atoi(ep + 1);
  const size_t z = frac.find_last_not_of('0');
  frac.erase(z == std::string::npos ? 0 : z + 1);
  if (e < 0 && e > -5) {
    frac = m + frac;
    m.clear();
    e++;
  }
  if (e > 0) {
    const int take = std::min<int>(e, (int)frac.size());
    m += frac.substr(0, take);
    frac.erase(0, take);
    e -= take;
  }
  while (e > -5 && e < 0 && m.empty()) {
    frac = "0" + frac;
    e++;
  }
  while (e < 3 && e > 0 && frac.empty()) {
    m += '0';
    e--;
  }
  std::string out = sign + m;
  if (!frac.empty()) out += "." + frac;
  if (e) out += "e" + std::to_string(e);
  return out;
}

static void dump_text(wb_tree *t, std::string &out)
{
  wb_ctx *c = t->ctx;
  const uint32_t nl = t->nleaves;
  std::vector<uint32_t> nf(nl), ni(nl);
  std::vector<double> cube(4 * (size_t)nl);
  if (nl) {
    uint32_t *d = (uint32_t *)wb_alloc(c, 8 * (size_t)nl);
    hipLaunchKernelGGL(k_leaf_dump_counts, dim3(nl), dim3(64), 0, c->stream, t->leaf_start,
                       t->leaf_cube, nl, t->px, t->py, t->pz, d, d + nl);
    WB_CHECK(hipMemcpyAsync(nf.data(), d, 4 * (size_t)nl, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipMemcpyAsync(ni.data(), d + nl, 4 * (size_t)nl, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipMemcpyAsync(cube.data(), t->leaf_cube, 32 * (size_t)nl, hipMemcpyDeviceToHost,
                            c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    wb_free(c, d, 8 * (size_t)nl);
  }
  uint64_t total = 0;
  for (uint32_t l = 0; l < nl; l++) {
    out += "(" + ldecimal(cube[4 * l], 0) + "," + ldecimal(cube[4 * l + 1], 0) + "," +
           ldecimal(cube[4 * l + 2], 0) + ")\xc2\xb1" + ldecimal(cube[4 * l + 3] / 2, 0) + " " +
           std::to_string(nf[l]) + " points";
    if (nf[l] != ni[l]) out += ", " + std::to_string((int)nf[l] - (int)ni[l]) + " stray points";
    out += "\n";
    total += nf[l];
  }
  out += std::to_string(total) + " total points\n";
}

extern "C" {

int wb_ldecimal(double x, double toler, char *buf, int cap)
{
  if (!buf || cap <= 0 || !(toler >= 0)) return WB_EINVAL;
  const std::string s = ldecimal(x, toler);
  if ((int)s.size() >= cap) return WB_ERANGE;
  memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

int wb_tree_dump(wb_tree *t, const char *path, char *buf, uint64_t cap, uint64_t *len)
{
  if (!t || (!path && !len)) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  try {
    WB_CHECK(hipSetDevice(c->device));
    std::string s;
    dump_text(t, s);
    if (len) *len = s.size();
    if (buf && cap) {
      const size_t m = std::min<size_t>(s.size(), cap - 1);
      memcpy(buf, s.data(), m);
      buf[m] = 0;
    }
    if (path) {
      FILE *fp = fopen(path, "wb");
      if (!fp) {
        c->err = std::string("cannot create ") + path;
        return WB_EINVAL;
      }
      const bool ok = fwrite(s.data(), 1, s.size(), fp) == s.size();
      if (fclose(fp) != 0 || !ok) {
        c->err = std::string("write failed: ") + path;
        return WB_EINVAL;
      }
    }
    return WB_OK;
  } catch (const std::string &m) {
    c->err = m;
    return WB_EDEVICE;
  }
}

}  // extern "C"
