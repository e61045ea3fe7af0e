// fv_mesh.cpp -- host-side mesh metric of the midpoint mesh, FP64.
//
// Computes, once per context, the arrays the reference's Mesh object holds
// (mesh/mesh_base.f90:558-672, mesh/mesh_midpoint.f90:97-145, scale factors
// mesh/geometry_logcylindrical.f90:310-342) and the static point-mass
// acceleration (sources/gravity_pointmass.f90:263-282 followed by the
// gravity list sum sources/sources_gravity.f90:226).  Operation order follows
// the reference so that the values are bit-identical; the device only ever
// sees the per-ring (i) and per-column (j) slices of these arrays.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "fv_host.h"

namespace fv {

static double hscale(const fv_config &c, double x) {
  return c.geometry == LOGCYLINDRICAL ? c.gparam * std::exp(x) : 1.0;
}

int mesh_build(const fv_config &c, Mesh &m) {
  if (c.geometry != CARTESIAN && c.geometry != LOGCYLINDRICAL) return -1;
  if (c.inum < 4 || c.jnum < 4) return -1;
  m.NI = c.inum + 2 * GH;
  m.NJ = c.jnum + 2 * GH;
  const int NI = m.NI, NJ = m.NJ;
  const size_t N = (size_t)NI * NJ;
  for (auto &a : m.a) a.assign(N, 0.0);
  m.dx = (c.xmax - c.xmin) / c.inum;
  m.dy = (c.ymax - c.ymin) / c.jnum;
  m.dz = (0.0 - 0.0) / 1;
  const double mdx = m.dx, mdy = m.dy;
  if (m.dz < 2 * DBL_EPSILON) m.dz = 1.0;  // knum = 1: mesh_dz = 0
  m.invdx = 1. / m.dx;
  m.invdy = 1. / m.dy;
  for (int j0 = 0; j0 < NJ; j0++) {
    const int fj = j0 - 1;  // Fortran index (JGMIN = -1)
    const double yc = c.ymin + ((double)(2 * fj - 1) * 0.5) * mdy;
    const double ys = yc + (-(0.5 * mdy)), yn = yc + 0.5 * mdy;
    for (int i0 = 0; i0 < NI; i0++) {
      const int fi = i0 - 1;
      const size_t k = (size_t)j0 * NI + i0;
      const double xc = c.xmin + ((double)(2 * fi - 1) * 0.5) * mdx;
      const double xw = xc + (-(0.5 * mdx)), xe = xc + 0.5 * mdx, xsn = xc + 0.0;
      const double hxc = hscale(c, xc), hyc = hxc, hzc = 1.0;
      const double hyW = hscale(c, xw), hyE = hscale(c, xe), hzW = 1.0, hzE = 1.0;
      const double hxS = hscale(c, xsn), hxN = hxS, hzS = 1.0, hzN = 1.0;
      m.a[M_BCX][k] = xc;
      m.a[M_BCY][k] = yc;
      m.a[M_DAXW][k] = hyW * hzW;
      m.a[M_DAXE][k] = hyE * hzE;
      m.a[M_DAYS][k] = hzS * hxS;
      m.a[M_DAYN][k] = hzN * hxN;
      const double sqrtg = hxc * (hyc * hzc);
      const double vol = sqrtg * m.dx * m.dy * m.dz;
      m.a[M_DXDYDV][k] = m.dx * m.dy / (vol + DBL_MIN);
      m.a[M_DYDZDV][k] = m.dy * m.dz / (vol + DBL_MIN);
      m.a[M_DZDXDV][k] = m.dz * m.dx / (vol + DBL_MIN);
      // commutator coefficients; bottom/top faces coincide with the centre
      m.a[M_CYXY][k] = 0.5 * (hzE + hzW) * (hyE - hyW) * m.a[M_DYDZDV][k];
      m.a[M_CXYX][k] = 0.5 * (hzN + hzS) * (hxN - hxS) * m.a[M_DZDXDV][k];
      m.a[M_CZXZ][k] = 0.5 * (hyE + hyW) * (hzE - hzW) * m.a[M_DYDZDV][k];
      m.a[M_CZYZ][k] = 0.5 * (hxN + hxS) * (hzN - hzS) * m.a[M_DZDXDV][k];
      // FARGO type 1 wfactor (timedisc_base.f90:428-447, direction 2): the
      // smallest hy over the W, E, bottom and top faces (bottom/top at the
      // centre's x) over hy at the bary-centre
      m.a[M_WFAC][k] = std::min(std::min(hyW, hyE), std::min(hxS, hxS)) / hyc;
      m.a[M_DLX][k] = hxc * m.dx;
      m.a[M_DLY][k] = hyc * m.dy;
      // reconstruction offsets faces - bary centre (reconstruction_linear.f90:156-177)
      m.a[M_RDXW][k] = xw - xc;
      m.a[M_RDXE][k] = xe - xc;
      m.a[M_RDYS][k] = ys - yc;
      m.a[M_RDYN][k] = yn - yc;
      if (c.geometry == LOGCYLINDRICAL) {
        const double r = c.gparam * std::exp(xc), rw = c.gparam * std::exp(xw);
        const double re = c.gparam * std::exp(xe), rs = c.gparam * std::exp(xsn);
        m.a[M_RAD][k] = std::sqrt(r * r + 0.0 * 0.0);
        m.a[M_RADW][k] = std::sqrt(rw * rw + 0.0 * 0.0);
        m.a[M_RADE][k] = std::sqrt(re * re + 0.0 * 0.0);
        m.a[M_RADS][k] = std::sqrt(rs * rs + 0.0 * 0.0);
        m.a[M_RADN][k] = m.a[M_RADS][k];
        m.a[M_POSX][k] = m.a[M_RAD][k];
        m.a[M_POSY][k] = 0.0;
      } else {
        m.a[M_RAD][k] = std::sqrt(xc * xc + yc * yc + 0.0 * 0.0);
        m.a[M_RADW][k] = std::sqrt(xw * xw + yc * yc);
        m.a[M_RADE][k] = std::sqrt(xe * xe + yc * yc);
        m.a[M_RADS][k] = std::sqrt(xc * xc + ys * ys);
        m.a[M_RADN][k] = std::sqrt(xc * xc + yn * yn);
        m.a[M_POSX][k] = xc;
        m.a[M_POSY][k] = yc;
      }
      if (c.gravity) {
        const double r = m.a[M_RAD][k];
        const double q = ((r * ((r - 0.0) * (r - 0.0))) + 0.0 * 0.0 * 0.0) / (c.GN * c.mass);
        const double om2 = 1. / (q > 10 * DBL_MIN ? q : 10 * DBL_MIN);
        m.a[M_ACCX][k] = 0.0 + (-(om2 * m.a[M_POSX][k]));
        m.a[M_ACCY][k] = 0.0 + (-(om2 * m.a[M_POSY][k]));
      }
    }
  }
  return 0;
}

}  // namespace fv

extern "C" int fv_mesh_array(const fv_config *cfg, int id, double *out) {
  if (!cfg || !out || id < 0 || id >= fv::M_COUNT) return -1;
  fv::Mesh m;
  if (fv::mesh_build(*cfg, m)) return -2;
  const auto &a = m.a[id];
  for (size_t k = 0; k < a.size(); k++) out[k] = a[k];
  return 0;
}
