// AMR hierarchy over batched levels (2D): coarse-fine ghost fill,
// flux-register reflux, conservative averaging and fine-first subcycling.
//
// The reference advances a hierarchy with SubcycleFineFirstSolver::
// AdvanceLevel (include/fub/solver/SubcycleFineFirstSolver.hpp:114-175):
// reset the coarse-fine register of the next finer level, advance that
// level `ratio` times with dt / ratio, advance this level
// (DimensionalSplitLevelIntegrator::AdvanceLevelNonRecursively,
// include/fub/solver/DimensionalSplitLevelIntegrator.hpp:221-276, with the
// split order of MakeSplitDirections, include/fub/split_method/
// SplittingMethod.hpp:115-126), reflux and average the finer level down.
// Ghost layers come from IntegratorContext::FillGhostLayerTwoLevels
// (src/AMReX/IntegratorContext.cpp:380-418: amrex::FillPatchTwoLevels with
// piecewise-constant interpolation) and the register from
// AccumulateCoarseFineFluxes / ApplyFluxCorrection / CoarsenConservatively
// (:451-533: amrex::FluxRegister, amrex::average_down).
//
// Here every copy, interpolation, register face and averaged cell is
// resolved once on the host into flat link lists when the hierarchy is
// created (the analogue of AMReX's cached copy plans); each operation is
// then one gather-style launch over its list, ordered on the handle's
// stream with the batched sweeps.  Patches of level l >= 1 must be aligned
// to the refinement ratio and properly nested: every ghost cell of a fine
// patch (clamped to the domain) lies in a fine patch or above a coarse
// patch's interior, and every coarse cell that receives a reflux
// correction belongs to a coarse patch.

#include "fvs_batch_internal.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

using namespace fvs;
using batch::kG;
using batch::PatchInfo;
using batch::RegFace;

namespace {

// Copy of all four components from one pool cell to another.
struct GhostLink {
  int64_t dst, src;
  int32_t dcs, scs;  // component strides of the destination and source patches
};

// One coarse cell covered by a fine patch: the mean of its ratio^2 fine cells.
struct AvgLink {
  int64_t dst, src;  // coarse cell; fine cell at the block's low corner
  int32_t dcs, scs, sldx, pad;
};

// One launch per ghost fill: links [0, nself) copy from the level's own
// pool, links [nself, nself + ncoarse) from the coarser level's pool 0.
// Every ghost cell appears in exactly one link and no link reads a ghost
// cell, so the copies are independent.
__global__ void GatherKernel(const GhostLink* self, int nself, const GhostLink* coarse,
                             int ncoarse, double* dst, const double* csrc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nself + ncoarse) return;
  const bool own = q < nself;
  const GhostLink L = own ? self[q] : coarse[q - nself];
  const double* src = own ? dst : csrc;
#pragma unroll
  for (int k = 0; k < 4; ++k) dst[L.dst + k * int64_t(L.dcs)] = src[L.src + k * int64_t(L.scs)];
}

// amrex::average_down for cell data: c = sum over the fine block (rows
// outer, x inner, starting from zero), coarse = (1 / ratio^2) * c.
__global__ void AvgDownKernel(const AvgLink* links, int n, double* coarse, const double* fine,
                              int ratio, double volfrac) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const AvgLink L = links[q];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double c = 0.0;
    for (int jj = 0; jj < ratio; ++jj)
      for (int ii = 0; ii < ratio; ++ii) c = c + fine[L.src + k * int64_t(L.scs) + jj * int64_t(L.sldx) + ii];
    coarse[L.dst + k * int64_t(L.dcs)] = volfrac * c;
  }
}

// Reflux: the coarse cell beside each coarse face takes -reg / V (cell left
// of / below the face) or +reg / V (right of / above), V the coarse cell
// volume.  One thread per corrected cell applies its entries in the order
// of the (direction, side) groups: (x, low), (x, high), (y, low), (y, high).
struct RefluxCell {
  int64_t at;        // coarse pool offset of the cell, component 0
  int32_t cs;        // its patch's component stride
  int32_t first, count;  // its entries in the ordered entry list
  double vol;
};

__global__ void RefluxKernel(const RefluxCell* cells, int ncells, const int32_t* entries,
                             const RegFace* faces, int nfaces, const double* reg, double* pool) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncells) return;
  const RefluxCell C = cells[q];
  double u[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) u[k] = pool[C.at + k * int64_t(C.cs)];
  for (int e = 0; e < C.count; ++e) {
    const int f = entries[C.first + e];
    const bool low = faces[f].side == 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double corr = DivRn(reg[k * int64_t(nfaces) + f], C.vol);
      u[k] = low ? u[k] - corr : u[k] + corr;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) pool[C.at + k * int64_t(C.cs)] = u[k];
}

// The levels' time steps of one coarse step: level l + 1 takes the level
// l step divided by the ratio (correctly rounded, as the host divides).
struct LevelDts {
  double* dst[8];
  int n;
  int ratio;
};

__global__ void LevelDtKernel(LevelDts d, const double* dtc) {
  if (threadIdx.x != 0) return;
  double dt = *dtc;
  for (int l = 0; l < d.n; ++l) {
    *d.dst[l] = dt;
    dt = DivRn(dt, static_cast<double>(d.ratio));
  }
}

__global__ void CopyKernel(const GhostLink* links, int n, double* dst, const double* src) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const GhostLink L = links[q];
#pragma unroll
  for (int k = 0; k < 4; ++k) dst[L.dst + k * int64_t(L.dcs)] = src[L.src + k * int64_t(L.scs)];
}

__global__ void FillKernel(double* p, int64_t n, double v) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < n) p[q] = v;
}

unsigned Blocks(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Density and pressure of a conservative cell, pressure through the array
// path's CompleteFromCons.
__device__ __forceinline__ void RhoP(const double* pool, int64_t at, int64_t cs,
                                     const Gas<double>& g, double* rho, double* p) {
  double q[6];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = pool[at + k * cs];
  (void)CompleteFromCons<2, double>(q, g);
  *rho = q[0];
  *p = q[4];
}

// ScalarGradientDetector::TagCellsForRefinement
// (include/fub/tagging_method/GradientDetector.hpp:150-194) with the
// projections density and pressure: a cell is tagged when, in either
// direction, a neighbour differs and |x_M - x_N| / (|x_M| + |x_N|) exceeds
// the projection's tolerance.  One thread per interior cell (blockIdx.y =
// patch); tags (one byte per interior cell, patches in order, x fastest)
// are overwritten.
__device__ __forceinline__ bool Jump(double xL, double xM, double xR, double tol) {
  if (!(xM != xR || xM != xL)) return false;
  const double left = DivRn(__builtin_fabs(xM - xL), __builtin_fabs(xM) + __builtin_fabs(xL));
  const double right = DivRn(__builtin_fabs(xM - xR), __builtin_fabs(xM) + __builtin_fabs(xR));
  return left > tol || right > tol;
}

__global__ void TagKernel(const PatchInfo* patches, const int64_t* tag_off, const double* pool,
                          Gas<double> g, double tol_rho, double tol_p, uint8_t* tags) {
  const PatchInfo P = patches[blockIdx.y];
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= int64_t(P.nx) * P.ny) return;
  const int64_t i = q % P.nx, j = q / P.nx;
  const int64_t ldx = P.nx + 2 * kG, cs = ldx * (P.ny + 2 * kG);
  const int64_t at = P.offset + (j + kG) * ldx + (i + kG);
  double r[5], p[5];  // centre, x - 1, x + 1, y - 1, y + 1
  const int64_t nb[5] = {at, at - 1, at + 1, at - ldx, at + ldx};
#pragma unroll
  for (int s = 0; s < 5; ++s) RhoP(pool, nb[s], cs, g, &r[s], &p[s]);
  bool tag = false;
  // directions in order, density before pressure (the condition order)
  tag |= Jump(r[1], r[0], r[2], tol_rho);
  tag |= Jump(p[1], p[0], p[2], tol_p);
  tag |= Jump(r[3], r[0], r[4], tol_rho);
  tag |= Jump(p[3], p[0], p[4], tol_p);
  tags[tag_off[blockIdx.y] + q] = tag ? 1 : 0;
}

} // namespace

struct fvs_amr {
  fvs_handle* h = nullptr;
  int nlevels = 0;
  int ratio = 2;
  int64_t n0[2] = {0, 0};
  std::vector<fvs_batch*> lv;
  struct Level {
    GhostLink* self = nullptr;  // sources in the same level
    int nself = 0;
    GhostLink* coarse = nullptr;  // sources in the next coarser level (pool 0)
    int ncoarse = 0;
    GhostLink* phys[2] = {nullptr, nullptr};  // ghost cells outside the domain
    int nphys[2] = {0, 0};                    // across x / y: the patch's edge row
    AvgLink* avg = nullptr;  // this level onto the next coarser one
    int navg = 0;
    RegFace* reg = nullptr;  // register between this level and the coarser one
    int nreg = 0;
    int group[5] = {0, 0, 0, 0, 0};  // entries of (dir 0, side 0), (0, 1), (1, 0), (1, 1)
    double* regval = nullptr;        // 4 x nreg
    RefluxCell* rcells = nullptr;    // coarse cells corrected by this register
    int nrcells = 0;
    int32_t* rentries = nullptr;     // their entries, per cell in group order
    int cfl_base = 0;                // first CFL slot of this level's sweeps
    int64_t* tag_off = nullptr;      // first tag of each patch
    int64_t ntags = 0;
    int64_t max_cells = 0;
  };
  std::vector<Level> L;
  int sequencing = FVS_AMR_SEQ_REFERENCE;  // ghost-layer sequencing of a level step
  std::vector<std::vector<int32_t>> lo;  // per level, (i0, j0) of every patch
  double* cfl = nullptr;  // one slot per sweep of a coarse step, then 2 per level
  int ncfl = 0;
  int64_t cycle = 0;
  double* dtc = nullptr;  // device: the coarse time step of the current advance
  unsigned long long* slot_partials = nullptr;  // per sweep of a step: pcap minima
  int pcap = 0;
  hipStream_t s = nullptr;  // stream the operations are enqueued on
  // whole coarse steps captured per cycle parity (the split order flips)
  int use_graphs = 1;
  hipStream_t cap = nullptr;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<int64_t> gens;  // the levels' tile-list generations the graphs were captured with
};

namespace {

template <class T>
int UploadVec(const std::vector<T>& v, T** dst) {
  *dst = nullptr;
  if (v.empty()) return FVS_OK;
  FVS_HIP_CHECK(hipMalloc(dst, v.size() * sizeof(T)));
  FVS_HIP_CHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return FVS_OK;
}

// Patch index covering each cell of a level (-1: none).
struct Owner {
  int64_t nx = 0, ny = 0;
  std::vector<int32_t> id;
  int32_t at(int64_t i, int64_t j) const {
    if (i < 0 || j < 0 || i >= nx || j >= ny) return -1;
    return id[j * nx + i];
  }
};

Owner OwnerOf(const fvs_amr* A, int l) {
  Owner O;
  int64_t scale = 1;
  for (int k = 0; k < l; ++k) scale *= A->ratio;
  O.nx = A->n0[0] * scale;
  O.ny = A->n0[1] * scale;
  O.id.assign(static_cast<size_t>(O.nx * O.ny), -1);
  const auto& hp = A->lv[l]->host_patches;
  for (size_t p = 0; p < hp.size(); ++p)
    for (int64_t j = A->lo[l][2 * p + 1]; j < A->lo[l][2 * p + 1] + hp[p].ny; ++j)
      for (int64_t i = A->lo[l][2 * p]; i < A->lo[l][2 * p] + hp[p].nx; ++i)
        O.id[j * O.nx + i] = static_cast<int32_t>(p);
  return O;
}

int64_t CellOffset(const PatchInfo& P, int64_t i, int64_t j) {
  return P.offset + (j + kG) * (P.nx + 2 * kG) + (i + kG);
}
int32_t CompStride(const PatchInfo& P) { return (P.nx + 2 * kG) * (P.ny + 2 * kG); }

int Build(fvs_amr* A, const int32_t* lo) {
  const int r = A->ratio;
  std::vector<Owner> own(A->nlevels);
  std::vector<std::vector<int32_t>> los(A->nlevels);
  int64_t base = 0;
  for (int l = 0; l < A->nlevels; ++l) {
    fvs_batch* b = A->lv[l];
    int64_t scale = 1;
    for (int k = 0; k < l; ++k) scale *= r;
    Owner& O = own[l];
    O.nx = A->n0[0] * scale;
    O.ny = A->n0[1] * scale;
    O.id.assign(static_cast<size_t>(O.nx * O.ny), -1);
    los[l].assign(lo + 2 * base, lo + 2 * (base + b->count));
    base += b->count;
    for (int64_t p = 0; p < b->count; ++p) {
      const PatchInfo& P = b->host_patches[p];
      const int64_t i0 = los[l][2 * p], j0 = los[l][2 * p + 1];
      if (i0 < 0 || j0 < 0 || i0 + P.nx > O.nx || j0 + P.ny > O.ny) {
        SetError("level %d patch %lld lies outside the level's domain", l, (long long)p);
        return FVS_BAD_ARGUMENT;
      }
      if (l > 0 && (i0 % r || j0 % r || P.nx % r || P.ny % r)) {
        SetError("level %d patch %lld is not aligned to the refinement ratio", l, (long long)p);
        return FVS_BAD_ARGUMENT;
      }
      for (int64_t j = j0; j < j0 + P.ny; ++j)
        for (int64_t i = i0; i < i0 + P.nx; ++i) {
          int32_t& o = O.id[j * O.nx + i];
          if (o >= 0) {
            SetError("level %d patches %d and %lld overlap", l, o, (long long)p);
            return FVS_BAD_ARGUMENT;
          }
          o = static_cast<int32_t>(p);
        }
    }
  }
  A->lo = los;
  for (int64_t c : own[0].id)
    if (c < 0) {
      SetError("level 0 patches must cover the domain");
      return FVS_BAD_ARGUMENT;
    }
  int slot = 0;
  for (int l = 0; l < A->nlevels; ++l) {
    fvs_batch* b = A->lv[l];
    const auto& hp = b->host_patches;
    const Owner& O = own[l];
    fvs_amr::Level& LV = A->L[l];
    int64_t sub = 1;
    for (int k = 0; k < l; ++k) sub *= r;
    LV.cfl_base = slot;
    slot += static_cast<int>(2 * sub);
    std::vector<GhostLink> self, coarse;
    // ghost layers: clamp to the domain (transmissive edges), then the
    // level's own patches, else the coarser level's cell above it
    for (int64_t p = 0; p < b->count; ++p) {
      const PatchInfo& P = hp[p];
      const int64_t i0 = los[l][2 * p], j0 = los[l][2 * p + 1];
      for (int64_t j = -kG; j < P.ny + kG; ++j)
        for (int64_t i = -kG; i < P.nx + kG; ++i) {
          if (i >= 0 && i < P.nx && j >= 0 && j < P.ny) continue;
          const int64_t I = std::clamp<int64_t>(i0 + i, 0, O.nx - 1);
          const int64_t J = std::clamp<int64_t>(j0 + j, 0, O.ny - 1);
          GhostLink g;
          g.dst = CellOffset(P, i, j);
          g.dcs = CompStride(P);
          const int32_t q = O.at(I, J);
          if (q >= 0) {
            const PatchInfo& Q = hp[q];
            g.src = CellOffset(Q, I - los[l][2 * q], J - los[l][2 * q + 1]);
            g.scs = CompStride(Q);
            self.push_back(g);
            continue;
          }
          const int32_t cq = l > 0 ? own[l - 1].at(I / r, J / r) : -1;
          if (cq < 0) {
            SetError("level %d patch %lld: ghost cell (%lld, %lld) has no source (patches "
                     "not properly nested)", l, (long long)p, (long long)(i0 + i),
                     (long long)(j0 + j));
            return FVS_BAD_ARGUMENT;
          }
          const PatchInfo& Q = A->lv[l - 1]->host_patches[cq];
          g.src = CellOffset(Q, I / r - los[l - 1][2 * cq], J / r - los[l - 1][2 * cq + 1]);
          g.scs = CompStride(Q);
          coarse.push_back(g);
        }
    }
    // TransmissiveBoundary of direction d (src/AMReX/boundary_condition/
    // TransmissiveBoundary.cpp:60-93): a ghost cell outside the domain
    // across d copies the patch's own cell in the domain's edge row, at the
    // same index along the other direction (ghost columns / rows included)
    std::vector<GhostLink> phys[2];
    for (int64_t p = 0; p < b->count; ++p) {
      const PatchInfo& P = hp[p];
      const int64_t lo_d[2] = {los[l][2 * p], los[l][2 * p + 1]};
      const int64_t n_d[2] = {P.nx, P.ny};
      const int64_t N_d[2] = {O.nx, O.ny};
      for (int d = 0; d < 2; ++d)
        for (int64_t k = -kG; k < n_d[d] + kG; ++k) {
          const int64_t K = lo_d[d] + k;
          if (K >= 0 && K < N_d[d]) continue;
          const int64_t edge = (K < 0 ? 0 : N_d[d] - 1) - lo_d[d];
          for (int64_t t = -kG; t < n_d[1 - d] + kG; ++t) {
            GhostLink g;
            g.dst = d == 0 ? CellOffset(P, k, t) : CellOffset(P, t, k);
            g.src = d == 0 ? CellOffset(P, edge, t) : CellOffset(P, t, edge);
            g.dcs = g.scs = CompStride(P);
            phys[d].push_back(g);
          }
        }
    }
    std::vector<AvgLink> avg;
    std::vector<RegFace> reg[4];
    if (l > 0) {
      const Owner& C = own[l - 1];
      const auto& cp = A->lv[l - 1]->host_patches;
      const auto& clo = los[l - 1];
      for (int64_t p = 0; p < b->count; ++p) {
        const PatchInfo& P = hp[p];
        const int64_t i0 = los[l][2 * p], j0 = los[l][2 * p + 1];
        for (int64_t cj = j0 / r; cj < (j0 + P.ny) / r; ++cj)
          for (int64_t ci = i0 / r; ci < (i0 + P.nx) / r; ++ci) {
            const int32_t q = C.at(ci, cj);
            if (q < 0) {
              SetError("level %d patch %lld is not above a level %d patch", l, (long long)p, l - 1);
              return FVS_BAD_ARGUMENT;
            }
            AvgLink a;
            a.dst = CellOffset(cp[q], ci - clo[2 * q], cj - clo[2 * q + 1]);
            a.dcs = CompStride(cp[q]);
            a.src = CellOffset(P, ci * r - i0, cj * r - j0);
            a.scs = CompStride(P);
            a.sldx = P.nx + 2 * kG;
            a.pad = 0;
            avg.push_back(a);
          }
        // register entries on the four sides
        for (int dir = 0; dir < 2; ++dir)
          for (int side = 0; side < 2; ++side) {
            const int64_t nd = dir == 0 ? P.nx : P.ny, nt = dir == 0 ? P.ny : P.nx;
            const int64_t d0 = dir == 0 ? i0 : j0, t0 = dir == 0 ? j0 : i0;
            const int64_t Nd = dir == 0 ? O.nx : O.ny;
            const int64_t face = side == 0 ? d0 : d0 + nd;  // fine face index along dir
            if (face == 0 || face == Nd) continue;         // domain edge
            const int64_t out = side == 0 ? face - 1 : face;  // fine cell outside the patch
            for (int64_t ct = t0 / r; ct < (t0 + nt) / r; ++ct) {
              const int64_t ft = ct * r;
              const int64_t oi = dir == 0 ? out : ft, oj = dir == 0 ? ft : out;
              if (O.at(oi, oj) >= 0) continue;  // fine-fine face
              const int64_t cface = face / r;
              const int64_t cc = side == 0 ? cface - 1 : cface;  // corrected coarse cell
              const int64_t cci = dir == 0 ? cc : ct, ccj = dir == 0 ? ct : cc;
              const int32_t q = C.at(cci, ccj);
              if (q < 0) {
                SetError("level %d patch %lld: coarse cell (%lld, %lld) beside it is in no "
                         "level %d patch", l, (long long)p, (long long)cci, (long long)ccj, l - 1);
                return FVS_BAD_ARGUMENT;
              }
              RegFace f;
              f.fp = static_cast<int32_t>(p);
              const int64_t fin = side == 0 ? 0 : nd;  // fine cell right of / above the face
              f.fi = static_cast<int32_t>(dir == 0 ? fin : ft - i0);
              f.fj = static_cast<int32_t>(dir == 0 ? ft - j0 : fin);
              f.cp = q;
              const int64_t cb_d = cface, cb_t = ct;  // coarse cell right of / above the face
              f.ci = static_cast<int32_t>((dir == 0 ? cb_d : cb_t) - clo[2 * q]);
              f.cj = static_cast<int32_t>((dir == 0 ? cb_t : cb_d) - clo[2 * q + 1]);
              f.side = side;
              f.pad = 0;
              reg[2 * dir + side].push_back(f);
            }
          }
      }
    }
    std::vector<RegFace> all;
    for (int g = 0; g < 4; ++g) {
      LV.group[g] = static_cast<int>(all.size());
      all.insert(all.end(), reg[g].begin(), reg[g].end());
    }
    LV.group[4] = static_cast<int>(all.size());
    std::vector<RefluxCell> rc;
    std::vector<int32_t> rent;
    if (l > 0) {
      // entries per corrected coarse cell, in entry (= group) order
      const auto& cp = A->lv[l - 1]->host_patches;
      std::vector<std::pair<int64_t, int32_t>> key;
      for (size_t e = 0; e < all.size(); ++e) {
        const RegFace& f = all[e];
        const int dir = static_cast<int>(e) < LV.group[2] ? 0 : 1;
        const int i = f.ci - (f.side == 0 && dir == 0 ? 1 : 0);
        const int j = f.cj - (f.side == 0 && dir == 1 ? 1 : 0);
        key.emplace_back(CellOffset(cp[f.cp], i, j), static_cast<int32_t>(e));
      }
      std::stable_sort(key.begin(), key.end(),
                       [](const auto& a, const auto& b) { return a.first < b.first; });
      for (size_t k = 0; k < key.size();) {
        size_t m = k;
        RefluxCell c;
        c.at = key[k].first;
        c.first = static_cast<int32_t>(rent.size());
        const RegFace& f = all[key[k].second];
        c.cs = CompStride(cp[f.cp]);
        c.vol = cp[f.cp].dx * cp[f.cp].dx;
        while (m < key.size() && key[m].first == key[k].first) rent.push_back(key[m++].second);
        c.count = static_cast<int32_t>(m - k);
        rc.push_back(c);
        k = m;
      }
    }
    std::vector<int64_t> toff(b->count);
    for (int64_t p = 0; p < b->count; ++p) {
      toff[p] = LV.ntags;
      LV.ntags += int64_t(hp[p].nx) * hp[p].ny;
      LV.max_cells = std::max<int64_t>(LV.max_cells, int64_t(hp[p].nx) * hp[p].ny);
    }
    if (int rc2 = UploadVec(toff, &LV.tag_off)) return rc2;
    LV.nrcells = static_cast<int>(rc.size());
    if (int rc2 = UploadVec(rc, &LV.rcells)) return rc2;
    if (int rc2 = UploadVec(rent, &LV.rentries)) return rc2;
    LV.nself = static_cast<int>(self.size());
    LV.ncoarse = static_cast<int>(coarse.size());
    LV.navg = static_cast<int>(avg.size());
    LV.nreg = static_cast<int>(all.size());
    if (int rc = UploadVec(self, &LV.self)) return rc;
    if (int rc = UploadVec(coarse, &LV.coarse)) return rc;
    for (int d = 0; d < 2; ++d) {
      LV.nphys[d] = static_cast<int>(phys[d].size());
      if (int rc = UploadVec(phys[d], &LV.phys[d])) return rc;
    }
    if (int rc = UploadVec(avg, &LV.avg)) return rc;
    if (int rc = UploadVec(all, &LV.reg)) return rc;
    if (LV.nreg) FVS_HIP_CHECK(hipMalloc(&LV.regval, 4 * sizeof(double) * LV.nreg));
  }
  A->ncfl = slot + 2 * A->nlevels;
  FVS_HIP_CHECK(hipMalloc(&A->cfl, A->ncfl * sizeof(double)));
  FVS_HIP_CHECK(hipMalloc(&A->dtc, sizeof(double)));
  A->pcap = A->lv[0]->cus * 8;  // the batched sweeps' largest grid
  FVS_HIP_CHECK(hipMalloc(&A->slot_partials, sizeof(unsigned long long) * A->pcap * slot));
  FVS_HIP_CHECK(hipStreamCreateWithFlags(&A->cap, hipStreamNonBlocking));
  return FVS_OK;
}

int FillGhosts(fvs_amr* A, int l, int which) {
  fvs_amr::Level& LV = A->L[l];
  fvs_batch* b = A->lv[l];
  const int n = LV.nself + LV.ncoarse;
  if (!n) return FVS_OK;
  hipLaunchKernelGGL(GatherKernel, dim3(Blocks(n)), dim3(256), 0, A->s, LV.self, LV.nself,
                     LV.coarse, LV.ncoarse, b->pool[which],
                     l > 0 ? A->lv[l - 1]->pool[0] : nullptr);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// The physical boundary of direction d on pool `which` of level l (the
// boundary condition the reference applies before every split direction
// but the first).
int FillPhysical(fvs_amr* A, int l, int which, int d) {
  fvs_amr::Level& LV = A->L[l];
  const int n = LV.nphys[d];
  if (!n) return FVS_OK;
  hipLaunchKernelGGL(GatherKernel, dim3(Blocks(n)), dim3(256), 0, A->s, LV.phys[d], n,
                     static_cast<const GhostLink*>(nullptr), 0, A->lv[l]->pool[which],
                     static_cast<const double*>(nullptr));
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int SetDt(fvs_amr* A, int l, double dt) {
  hipLaunchKernelGGL(FillKernel, dim3(1), dim3(64), 0, A->s, A->lv[l]->dts, int64_t(1), dt);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// Arguments of the register contributions of register f (between level f
// and level f - 1) from a sweep in dir of the fine (fine = 1) or the coarse
// level reading pool src; count 0 when the register has no such entries.
batch::RegArgs RegisterArgs(fvs_amr* A, int f, int dir, int fine, int src) {
  fvs_amr::Level& LV = A->L[f];
  const int first = LV.group[2 * dir], count = LV.group[2 * dir + 2] - first;
  batch::RegArgs a{};
  if (count == 0) return a;
  fvs_batch* b = A->lv[fine ? f : f - 1];
  a.faces = LV.reg;
  a.first = first;
  a.count = count;
  a.nfaces = LV.nreg;
  a.patches = b->patches;
  a.pool = b->pool[src];
  a.dt = b->dts;
  a.reg = LV.regval;
  a.g.gamma = A->h->desc.gamma;
  a.g.gm1inv = 1.0 / (A->h->desc.gamma - 1.0);
  a.g.gm1 = A->h->desc.gamma - 1.0;
  a.limiter = A->h->desc.limiter;
  a.ratio = fine ? A->ratio : 1;
  a.fine = fine;
  a.sign = fine ? 1.0 : -1.0;
  return a;
}

int Reflux(fvs_amr* A, int f) {
  fvs_amr::Level& LV = A->L[f];
  if (!LV.nrcells) return FVS_OK;
  hipLaunchKernelGGL(RefluxKernel, dim3(Blocks(LV.nrcells)), dim3(256), 0, A->s, LV.rcells,
                     LV.nrcells, LV.rentries, LV.reg, LV.nreg, LV.regval, A->lv[f - 1]->pool[0]);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int AverageDown(fvs_amr* A, int f) {
  fvs_amr::Level& LV = A->L[f];
  if (!LV.navg) return FVS_OK;
  const double volfrac = 1.0 / static_cast<double>(A->ratio * A->ratio);
  hipLaunchKernelGGL(AvgDownKernel, dim3(Blocks(LV.navg)), dim3(256), 0, A->s, LV.avg,
                     LV.navg, A->lv[f - 1]->pool[0], A->lv[f]->pool[0], A->ratio, volfrac);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// SubcycleFineFirstSolver::AdvanceLevel with the level's split step; the
// level's state is in pool 0 before and after.  sub = (index, count) of
// this level's subcycle within its parent's step; slot = this level's
// next CFL slot.
int AdvanceLevel(fvs_amr* A, int l, int sub_first, std::vector<int>& slot) {
  if (int rc = FillGhosts(A, l, 0)) return rc;
  const bool finer = l + 1 < A->nlevels;
  if (finer) {
    fvs_amr::Level& F = A->L[l + 1];
    if (F.nreg) {
      hipLaunchKernelGGL(FillKernel, dim3(Blocks(4 * int64_t(F.nreg))), dim3(256), 0, A->s,
                         F.regval, 4 * int64_t(F.nreg), 0.0);
      FVS_LAUNCH_CHECK();
    }
    for (int s = 0; s < A->ratio; ++s) {
      if (int rc = AdvanceLevel(A, l + 1, s, slot)) return rc;
    }
  }
  // MakeSplitDirections: reversed order when (cycle % 2 + subcycle) % 2 == 1
  const bool rev = ((A->cycle % 2) + sub_first) % 2 == 1;
  const int dirs[2] = {rev ? 1 : 0, rev ? 0 : 1};
  // Reference sequencing: the ghost layers filled above serve the whole
  // level step; the first sweep also advances the ghost rows across its
  // direction, and the second refills only its direction's physical
  // boundary.  Refill sequencing: every sweep refills all ghost layers.
  const bool ref = A->sequencing == FVS_AMR_SEQ_REFERENCE;
  int src = 0;
  for (int w = 0; w < 2; ++w) {
    const int d = dirs[w];
    if (w > 0) {
      if (int rc = ref ? FillPhysical(A, l, src, d) : FillGhosts(A, l, src)) return rc;
    }
    // the sweep's register contributions: to the finer level's register (as
    // its coarse side) and to this level's own (as its fine side), one launch
    {
      const batch::RegArgs rc0 = finer ? RegisterArgs(A, l + 1, d, 0, src) : batch::RegArgs{};
      const batch::RegArgs rc1 = l > 0 ? RegisterArgs(A, l, d, 1, src) : batch::RegArgs{};
      batch::LaunchRegFlux(A->s, d, rc0, rc1, A->h->desc.precision);
      FVS_LAUNCH_CHECK();
    }
    const bool grown = ref && w == 0;
    if ((grown ? A->lv[l]->grid_grown[d] : A->lv[l]->grid[d]) > static_cast<unsigned>(A->pcap)) {
      SetError("batched sweep grid exceeds the per-sweep minima buffer");
      return FVS_BAD_ARGUMENT;
    }
    const int q = slot[l]++;
    if (int rc = batch::SweepPartials(A->lv[l], A->s, d, src,
                                      A->slot_partials + static_cast<int64_t>(q) * A->pcap,
                                      grown))
      return rc;
    src = 1 - src;
  }
  if (finer) {
    if (int rc = Reflux(A, l + 1)) return rc;
    if (int rc = AverageDown(A, l + 1)) return rc;
  }
  return FVS_OK;
}

// One coarse step on stream A->s: the levels' time steps from *A->dtc, then
// the recursive level steps.
int EnqueueStep(fvs_amr* A) {
  LevelDts d;
  d.n = A->nlevels;
  d.ratio = A->ratio;
  for (int l = 0; l < A->nlevels; ++l) d.dst[l] = A->lv[l]->dts;
  hipLaunchKernelGGL(LevelDtKernel, dim3(1), dim3(64), 0, A->s, d, A->dtc);
  FVS_LAUNCH_CHECK();
  // minima of sweeps with smaller grids must not see an earlier step's
  const int nsweeps = A->ncfl - 2 * A->nlevels;
  if (int rc = batch::LaunchSlotMin(A->s, A->slot_partials, nsweeps, A->pcap, nullptr, true))
    return rc;
  std::vector<int> slot(A->nlevels);
  for (int l = 0; l < A->nlevels; ++l) slot[l] = A->L[l].cfl_base;
  if (int rc = AdvanceLevel(A, 0, 0, slot)) return rc;
  return batch::LaunchSlotMin(A->s, A->slot_partials, nsweeps, A->pcap, A->cfl, false);
}

} // namespace

extern "C" {

int fvs_amr_create(fvs_handle* h, int nlevels, fvs_batch* const* levels, const int32_t* lo,
                   const int64_t* domain, int ratio, fvs_amr** out) {
  if (!h || nlevels < 1 || nlevels > 8 || !levels || !lo || !domain || !out || ratio < 2) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  *out = nullptr;
  if (domain[0] < 1 || domain[1] < 1) {
    SetError("bad domain size");
    return FVS_BAD_ARGUMENT;
  }
  fvs_amr* A = new (std::nothrow) fvs_amr();
  if (!A) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  A->h = h;
  A->nlevels = nlevels;
  A->ratio = ratio;
  A->n0[0] = domain[0];
  A->n0[1] = domain[1];
  A->lv.assign(levels, levels + nlevels);
  A->L.resize(nlevels);
  for (int l = 0; l < nlevels; ++l) {
    if (!levels[l] || levels[l]->h != h) {
      delete A;
      SetError("level %d: missing batch or batch of another handle", l);
      return FVS_BAD_ARGUMENT;
    }
  }
  if (int rc = Build(A, lo)) {
    fvs_amr_destroy(A);
    return rc;
  }
  *out = A;
  return FVS_OK;
}

int fvs_amr_destroy(fvs_amr* A) {
  if (!A) return FVS_OK;
  if (A->h) (void)hipStreamSynchronize(A->h->stream);
  if (A->cap) (void)hipStreamSynchronize(A->cap);
  for (hipGraphExec_t& e : A->exec)
    if (e) (void)hipGraphExecDestroy(e);
  if (A->cap) (void)hipStreamDestroy(A->cap);
  if (A->dtc) (void)hipFree(A->dtc);
  if (A->slot_partials) (void)hipFree(A->slot_partials);
  for (auto& LV : A->L) {
    void* ps[] = {LV.self, LV.coarse, LV.phys[0], LV.phys[1], LV.avg, LV.reg, LV.regval,
                  LV.rcells, LV.rentries, LV.tag_off};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  if (A->cfl) (void)hipFree(A->cfl);
  delete A;
  return FVS_OK;
}

int fvs_amr_fill_ghosts(fvs_amr* A, int level, int which) {
  if (!A || level < 0 || level >= A->nlevels || (which != 0 && which != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  A->s = A->h->stream;
  return FillGhosts(A, level, which);
}

int fvs_amr_average_down(fvs_amr* A, int fine_level) {
  if (!A || fine_level < 1 || fine_level >= A->nlevels) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  A->s = A->h->stream;
  return AverageDown(A, fine_level);
}

int fvs_amr_advance(fvs_amr* A, double dt) {
  if (!A || !(dt > 0.0)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  hipStream_t hs = A->h->stream;
  // the step's dt goes to device memory as a kernel argument (no host
  // buffer has to outlive the call); captured steps read it from there
  hipLaunchKernelGGL(FillKernel, dim3(1), dim3(64), 0, hs, A->dtc, int64_t(1), dt);
  FVS_LAUNCH_CHECK();
  if (!A->use_graphs) {
    A->s = hs;
    if (int rc = EnqueueStep(A)) return rc;
  } else {
    // a level whose tile lists changed (autotuning, a new shape) invalidates
    // the captured steps, which hold its old lists
    std::vector<int64_t> gens(A->nlevels);
    for (int l = 0; l < A->nlevels; ++l) gens[l] = A->lv[l]->gen;
    if (gens != A->gens) {
      FVS_HIP_CHECK(hipStreamSynchronize(hs));
      for (hipGraphExec_t& e : A->exec)
        if (e) {
          (void)hipGraphExecDestroy(e);
          e = nullptr;
        }
      A->gens = gens;
    }
    hipGraphExec_t& exec = A->exec[A->cycle % 2];
    if (!exec) {
      // capture on the hierarchy's own stream (the handle's may be the null
      // stream, which cannot capture); nothing runs while capturing
      A->s = A->cap;
      FVS_HIP_CHECK(hipStreamBeginCapture(A->cap, hipStreamCaptureModeThreadLocal));
      const int rc = EnqueueStep(A);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(A->cap, &g);
      A->s = hs;
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      FVS_HIP_CHECK(e);
      const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      FVS_HIP_CHECK(ie);
    }
    FVS_HIP_CHECK(hipGraphLaunch(exec, hs));
  }
  A->cycle += 1;
  return FVS_OK;
}

int fvs_amr_tag_cells(fvs_amr* A, int level, double tol_rho, double tol_p, uint8_t* tags) {
  if (!A || level < 0 || level >= A->nlevels || !tags || !(tol_rho >= 0.0) || !(tol_p >= 0.0)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  A->s = A->h->stream;
  if (int rc = FillGhosts(A, level, 0)) return rc;
  fvs_batch* b = A->lv[level];
  const fvs_amr::Level& LV = A->L[level];
  if (b->count > 65535) {
    SetError("tagging: more than 65535 patches in a level");
    return FVS_BAD_ARGUMENT;
  }
  Gas<double> g;
  g.gamma = A->h->desc.gamma;
  g.gm1inv = 1.0 / (A->h->desc.gamma - 1.0);
  g.gm1 = A->h->desc.gamma - 1.0;
  hipLaunchKernelGGL(TagKernel, dim3(Blocks(LV.max_cells), static_cast<unsigned>(b->count)),
                     dim3(256), 0, A->s, b->patches, LV.tag_off, b->pool[0], g, tol_rho, tol_p,
                     tags);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_amr_transfer(fvs_amr* D, fvs_amr* S) {
  if (!D || !S || D == S || D->h != S->h || D->nlevels != S->nlevels || D->ratio != S->ratio ||
      D->n0[0] != S->n0[0] || D->n0[1] != S->n0[1]) {
    SetError("transfer needs two hierarchies of one handle with the same domain, ratio and "
             "number of levels");
    return FVS_BAD_ARGUMENT;
  }
  const int r = D->ratio;
  hipStream_t s = D->h->stream;
  for (int l = 0; l < D->nlevels; ++l) {
    const Owner old = OwnerOf(S, l);
    const Owner crs = l > 0 ? OwnerOf(D, l - 1) : Owner();
    const auto& dp = D->lv[l]->host_patches;
    const auto& sp = S->lv[l]->host_patches;
    std::vector<GhostLink> from_old, from_coarse;
    for (size_t p = 0; p < dp.size(); ++p) {
      const int64_t i0 = D->lo[l][2 * p], j0 = D->lo[l][2 * p + 1];
      for (int64_t j = 0; j < dp[p].ny; ++j)
        for (int64_t i = 0; i < dp[p].nx; ++i) {
          const int64_t I = i0 + i, J = j0 + j;
          GhostLink g;
          g.dst = CellOffset(dp[p], i, j);
          g.dcs = CompStride(dp[p]);
          const int32_t q = old.at(I, J);
          if (q >= 0) {
            g.src = CellOffset(sp[q], I - S->lo[l][2 * q], J - S->lo[l][2 * q + 1]);
            g.scs = CompStride(sp[q]);
            from_old.push_back(g);
            continue;
          }
          const int32_t cq = l > 0 ? crs.at(I / r, J / r) : -1;
          if (cq < 0) {
            SetError("transfer: level %d cell (%lld, %lld) has no source", l, (long long)I,
                     (long long)J);
            return FVS_BAD_ARGUMENT;
          }
          const PatchInfo& C = D->lv[l - 1]->host_patches[cq];
          g.src = CellOffset(C, I / r - D->lo[l - 1][2 * cq], J / r - D->lo[l - 1][2 * cq + 1]);
          g.scs = CompStride(C);
          from_coarse.push_back(g);
        }
    }
    GhostLink *a = nullptr, *b = nullptr;
    int rc = UploadVec(from_old, &a);
    if (!rc) rc = UploadVec(from_coarse, &b);
    if (!rc && !from_old.empty())
      hipLaunchKernelGGL(CopyKernel, dim3(Blocks(from_old.size())), dim3(256), 0, s, a,
                         static_cast<int>(from_old.size()), D->lv[l]->pool[0], S->lv[l]->pool[0]);
    if (!rc && !from_coarse.empty())
      hipLaunchKernelGGL(CopyKernel, dim3(Blocks(from_coarse.size())), dim3(256), 0, s, b,
                         static_cast<int>(from_coarse.size()), D->lv[l]->pool[0],
                         D->lv[l - 1]->pool[0]);
    if (!rc && hipGetLastError() != hipSuccess) rc = FVS_HIP_ERROR;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = FVS_HIP_ERROR;
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (rc) {
      if (rc == FVS_HIP_ERROR) SetError("transfer: HIP error");
      return rc;
    }
  }
  D->cycle = S->cycle;
  return FVS_OK;
}

int fvs_amr_cycle(fvs_amr* A, int64_t* cycle) {
  if (!A || !cycle) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  *cycle = A->cycle;
  return FVS_OK;
}

int fvs_amr_set_cycle(fvs_amr* A, int64_t cycle) {
  if (!A || cycle < 0) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  A->cycle = cycle;
  return FVS_OK;
}

int fvs_amr_set_sequencing(fvs_amr* A, int mode) {
  if (!A || (mode != FVS_AMR_SEQ_REFERENCE && mode != FVS_AMR_SEQ_REFILL)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  if (mode != A->sequencing) {
    // the captured steps hold the other sequence of launches
    FVS_HIP_CHECK(hipStreamSynchronize(A->h->stream));
    for (hipGraphExec_t& e : A->exec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
    A->sequencing = mode;
  }
  return FVS_OK;
}

int fvs_amr_set_graphs(fvs_amr* A, int on) {
  if (!A) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  A->use_graphs = on ? 1 : 0;
  return FVS_OK;
}

int fvs_amr_level_cfl(fvs_amr* A, double* host) {
  if (!A || !host) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  std::vector<double> v(A->ncfl);
  FVS_HIP_CHECK(hipMemcpyAsync(v.data(), A->cfl, A->ncfl * sizeof(double),
                               hipMemcpyDeviceToHost, A->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(A->h->stream));
  for (int l = 0; l < A->nlevels; ++l) {
    const int n = (l + 1 < A->nlevels ? A->L[l + 1].cfl_base : A->ncfl - 2 * A->nlevels) -
                  A->L[l].cfl_base;
    double m = __builtin_huge_val();
    for (int k = 0; k < n; ++k) m = std::min(m, v[A->L[l].cfl_base + k]);
    host[l] = m;
  }
  return FVS_OK;
}

int fvs_amr_stable_dt(fvs_amr* A, double* dt_host) {
  if (!A || !dt_host) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  const int base = A->ncfl - 2 * A->nlevels;
  A->s = A->h->stream;
  for (int l = 0; l < A->nlevels; ++l) {
    if (int rc = FillGhosts(A, l, 0)) return rc;
    if (int rc = SetDt(A, l, 0.0)) return rc;
    // sweeps into the scratch pool 1 for their CFL minima only
    for (int d = 0; d < 2; ++d)
      if (int rc = batch::Sweep(A->lv[l], A->s, d, 0, A->cfl + base + 2 * l + d)) return rc;
  }
  std::vector<double> v(2 * A->nlevels);
  FVS_HIP_CHECK(hipMemcpyAsync(v.data(), A->cfl + base, v.size() * sizeof(double),
                               hipMemcpyDeviceToHost, A->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(A->h->stream));
  double m = __builtin_huge_val();
  double scale = 1.0;
  for (int l = 0; l < A->nlevels; ++l) {
    m = std::min(m, scale * std::min(v[2 * l], v[2 * l + 1]));
    scale *= A->ratio;
  }
  *dt_host = m;
  return FVS_OK;
}

} // extern "C"
