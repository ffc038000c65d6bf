// hdg_ctx.h -- the host-side context of the per-cell HDG path, shared by the
// C ABI (hdg_capi.hip) and the trace solver (hdg_trace_solver.hip).
#pragma once
#include <string>
#include <vector>

#include "../../include/ddhdg_hip.h"
#include "hdg_kernels.h"
#include "hdg_tables.h"
#include "hdg_topology.h"

struct hdg_ctx
{
  hdg_mesh_desc mesh;
  hdg_phys      phys;
  int           d, k, nb, nf, qc, qf, nfc, NC;
  long long     ncells, nfaces;
  int           n[3];
  double        h[3], lo[3];
  long long     face_off[4], face_dims[3][3];
  int           en[3], ne, encomp[3];
  int           variant = 0; // 0 auto, 1 generic, 2 register-resident
  int           precision = 0; // 0 FP64, 1 FP32 elimination + FP64 refinement
  hdg::HostTables ht;
  hdg::Topology   topo;  // numbering and trace block pattern
  std::vector<int> color_cells[2];
  // device
  double         *d_tab = nullptr;
  hdg::DevTables  tb{};
  double         *d_doping = nullptr, *d_dirv = nullptr, *d_neuv = nullptr;
  double         *d_taucf  = nullptr; // per (cell, face, comp) tau, or null
  double         *d_rgscr  = nullptr; // the generic kernel's global scratch (3D degree 2)
  long long      *d_blkptr = nullptr;
  int            *d_nbk    = nullptr;
  unsigned char  *d_cslot  = nullptr;
  int            *d_cells[2] = {nullptr, nullptr};
};

namespace hdg
{
  // sets the thread-local message hdg_last_error() returns; returns code
  int set_error(const std::string &msg, int code);

  inline long long cell_face(const hdg_ctx *C, long long c, int lf)
  {
    return C->topo.cell_face(c, lf);
  }
  inline void cell_coords(const hdg_ctx *C, long long c, int ic[3])
  {
    C->topo.cell_coords(c, ic);
  }
  inline int boundary_id(const hdg_ctx *C, long long c, int lf)
  {
    return C->topo.boundary_id(c, lf);
  }
} // namespace hdg
