// fv_host.h -- host-side types of libfosite_fv.
#pragma once
#include <string>
#include <vector>

#include "../../include/fosite_fv.h"
#include "fv_device.h"

namespace fv {

// mesh array ids, in the order documented for fv_mesh_array()
enum MeshId {
  M_DAXW, M_DAXE, M_DAYS, M_DAYN, M_DYDZDV, M_DZDXDV, M_DXDYDV, M_DLX, M_DLY,
  M_CXYX, M_CYXY, M_CZXZ, M_CZYZ, M_RDXW, M_RDXE, M_RDYS, M_RDYN, M_RAD,
  M_RADW, M_RADE, M_RADS, M_RADN, M_POSX, M_POSY, M_BCX, M_BCY, M_ACCX, M_ACCY,
  M_WFAC, M_COUNT
};

struct Mesh {
  int NI = 0, NJ = 0;
  double dx = 0, dy = 0, dz = 0, invdx = 0, invdy = 0;
  std::vector<double> a[M_COUNT];  // [NJ][NI] each
};

int mesh_build(const fv_config &c, Mesh &m);

}  // namespace fv
