// nb_lists.hip -- spatial sort and Verlet neighbor list on the GPU
// (NeighborList.java:351-372, 514-578, 777-1120, 1296-1545 restated for a
// full list in spatially sorted order, built 16 lanes per atom).
//
// Atoms are binned by fractional coordinate into cells at least one list
// cutoff wide, counting-sorted by cell (ties broken by caller index, so the
// order is deterministic), and each sorted atom scans the 27 (or fewer for
// thin boxes) neighbouring cells.  The list is full (both directions) so the
// pair kernels accumulate only their own atom and need no atomics.  The list
// is rebuilt only when some atom moved more than half the buffer since the
// last build (NeighborList.java:314, 777-802).
#include <stdexcept>

#include "nb_kern.h"

namespace {

__global__ void k_max_disp(int n, const double *__restrict__ xyz, const double *__restrict__ xyz0, double lim2,
                           int32_t *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dx = xyz[3 * i] - xyz0[3 * i], dy = xyz[3 * i + 1] - xyz0[3 * i + 1], dz = xyz[3 * i + 2] - xyz0[3 * i + 2];
  if (dx * dx + dy * dy + dz * dz > lim2) atomicOr(&flags[3], 1);
}

__device__ __forceinline__ int wrapi(int i, int n) {
  int m = i % n;
  return m < 0 ? m + n : m;
}

__device__ __forceinline__ int cell_of(const CellD &c, int nc0, int nc1, int nc2, double x, double y, double z) {
  int ci[3];
  double f[3];
  if (c.aperiodic) return 0;
  f[0] = x * c.A00 + y * c.A10 + z * c.A20;
  f[1] = y * c.A11 + z * c.A21;
  f[2] = z * c.A22;
  int nc[3] = {nc0, nc1, nc2};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double w = f[d] - floor(f[d]);
    int q = (int)(w * nc[d]);
    ci[d] = q >= nc[d] ? nc[d] - 1 : (q < 0 ? 0 : q);
  }
  return ci[0] + nc0 * (ci[1] + nc1 * ci[2]);
}

__global__ void k_cell_assign(int n, const double *__restrict__ xyz, CellD cell, int nc0, int nc1, int nc2,
                              int32_t *__restrict__ cellid, int32_t *__restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of(cell, nc0, nc1, nc2, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  cellid[i] = c;
  atomicAdd(&count[c], 1);
}

// exclusive scan of count[0..m) into start[0..m], single workgroup
__global__ void k_scan(int m, const int32_t *__restrict__ count, int32_t *__restrict__ start) {
  __shared__ int32_t part[1024];
  int t = threadIdx.x, nt = blockDim.x;
  int chunk = (m + nt - 1) / nt;
  int b = t * chunk, e = min(m, b + chunk);
  int s = 0;
  for (int i = b; i < e; i++) s += count[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int i = 0; i < nt; i++) {
      int v = part[i];
      part[i] = acc;
      acc += v;
    }
    start[m] = acc;
  }
  __syncthreads();
  int acc = part[t];
  for (int i = b; i < e; i++) {
    start[i] = acc;
    acc += count[i];
  }
}

__global__ void k_cell_scatter(int n, const int32_t *__restrict__ cellid, const int32_t *__restrict__ start,
                               int32_t *__restrict__ cursor, int32_t *__restrict__ perm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cellid[i];
  int slot = atomicAdd(&cursor[c], 1);
  perm[start[c] + slot] = i;
}

// Order each cell's segment by an 8x8x8 Morton key of the atom's place inside
// the cell, then by caller index (deterministic): atoms that follow each other
// in the sorted order are close in space, so the rows of the atoms one
// workgroup handles share most of their partners in the caches.  One
// workgroup per cell ranks every member by counting the smaller keys.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 7u;
  return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4);
}

__global__ void __launch_bounds__(256) k_cell_rank(int m, const int32_t *__restrict__ start,
                                                   const int32_t *__restrict__ unsorted, int32_t *__restrict__ perm,
                                                   const double *__restrict__ xyz, CellD cell, int nc0, int nc1,
                                                   int nc2) {
  __shared__ unsigned long long mem[2048];
  int c = blockIdx.x;
  if (c >= m) return;
  int b = start[c], e = start[c + 1], cnt = e - b;
  bool in_lds = cnt <= 2048;
  auto key_of = [&](int v) -> unsigned long long {
    uint32_t mc = 0;
    if (!cell.aperiodic) {
      double x = xyz[3 * v], y = xyz[3 * v + 1], z = xyz[3 * v + 2];
      double f[3] = {x * cell.A00 + y * cell.A10 + z * cell.A20, y * cell.A11 + z * cell.A21, z * cell.A22};
      int ncs[3] = {nc0, nc1, nc2};
      uint32_t sub[3];
      for (int d = 0; d < 3; d++) {
        double w = (f[d] - floor(f[d])) * ncs[d];
        w -= floor(w);
        int q = (int)(w * 8.0);
        sub[d] = (uint32_t)(q < 0 ? 0 : (q > 7 ? 7 : q));
      }
      mc = spread3(sub[0]) | (spread3(sub[1]) << 1) | (spread3(sub[2]) << 2);
    }
    return ((unsigned long long)mc << 32) | (uint32_t)v;
  };
  if (in_lds)
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) mem[t] = key_of(unsorted[b + t]);
  __syncthreads();
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    unsigned long long kv = in_lds ? mem[t] : key_of(unsorted[b + t]);
    int rank = 0;
    for (int j = 0; j < cnt; j++) rank += (in_lds ? mem[j] : key_of(unsorted[b + j])) < kv;
    perm[b + rank] = (int32_t)(kv & 0xffffffffull);
  }
}

__global__ void k_iperm(int n, const int32_t *__restrict__ perm, int32_t *__restrict__ iperm) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) iperm[perm[s]] = s;
}

// static per-atom data in sorted order
__global__ void k_gather_static(int n, const int32_t *__restrict__ perm, const double *__restrict__ polar0,
                                const double *__restrict__ thole0, const double *__restrict__ ipd0,
                                const int32_t *__restrict__ group0, const int32_t *__restrict__ vclass0,
                                const int32_t *__restrict__ soft0, const int32_t *__restrict__ mol0,
                                double *__restrict__ polar, int32_t *__restrict__ group, int32_t *__restrict__ vclass,
                                int32_t *__restrict__ soft, int32_t *__restrict__ mol, AtomMP *__restrict__ mp) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  polar[s] = polar0[i];
  group[s] = group0[i];
  vclass[s] = vclass0[i];
  if (soft0) {
    soft[s] = soft0[i] != 0;
    mol[s] = mol0[i];
  }
  mp[s].ipd = ipd0[i];
  mp[s].thole = thole0[i];
}

// LPA lanes per sorted atom scan the neighbouring cells LPA candidates at a
// time; hits are compacted in candidate order with a ballot, so each row lists
// its neighbours in (cell, sorted index) order exactly as a serial scan would.
template <int LPA>
__global__ void __launch_bounds__(256) k_build_list(int n, const double4 *__restrict__ pos, const int32_t *__restrict__ perm,
                                                    const int32_t *__restrict__ start, CellD cell, int nc0, int nc1,
                                                    int nc2, double cut2, const int32_t *__restrict__ group,
                                                    const int32_t *__restrict__ moff, const int32_t *__restrict__ matom,
                                                    const int32_t *__restrict__ mrel, uint32_t *__restrict__ nbr,
                                                    int maxnb, int32_t *__restrict__ nnb, int32_t *__restrict__ flags) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  if (s >= n) return;  // whole groups leave together
  double4 pi = pos[s];
  int ci = cell_of(cell, nc0, nc1, nc2, pi.x, pi.y, pi.z);
  int cx = ci % nc0, cy = (ci / nc0) % nc1, cz = ci / (nc0 * nc1);
  int lo0 = nc0 >= 3 ? -1 : 0, hi0 = nc0 >= 2 ? 1 : 0;
  int lo1 = nc1 >= 3 ? -1 : 0, hi1 = nc1 >= 2 ? 1 : 0;
  int lo2 = nc2 >= 3 ? -1 : 0, hi2 = nc2 >= 2 ? 1 : 0;
  int i = perm[s];
  int gi = group[s];
  int cnt = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int gbase = (threadIdx.x & 63) & ~(LPA - 1);
  uint32_t *row = nbr + (size_t)s * maxnb;
  for (int dz = lo2; dz <= hi2; dz++)
    for (int dy = lo1; dy <= hi1; dy++)
      for (int dx = lo0; dx <= hi0; dx++) {
        int q = wrapi(cx + dx, nc0) + nc0 * (wrapi(cy + dy, nc1) + nc1 * wrapi(cz + dz, nc2));
        int b = start[q], e = start[q + 1];
        for (int base = b; base < e; base += LPA) {
          int t = base + lane;
          bool hit = false;
          if (t < e && t != s) {
            double4 pk = pos[t];
            double ddx = pk.x - pi.x, ddy = pk.y - pi.y, ddz = pk.z - pi.z;
            hit = image(cell, ddx, ddy, ddz) <= cut2;
          }
          unsigned long long m = (__ballot(hit) >> gbase) & ((LPA == 64) ? ~0ull : ((1ull << LPA) - 1ull));
          if (hit) {
            int slot = cnt + __popcll(m & below);
            int code = relation(moff, matom, mrel, i, perm[t]) + (group[t] == gi ? 8 : 0);
            if (slot < maxnb) row[slot] = (uint32_t)t | ((uint32_t)code << NB_CODE_SHIFT);
          }
          cnt += __popcll(m);
        }
      }
  if (lane == 0) {
    nnb[s] = cnt;
    if (cnt > maxnb) atomicMax(&flags[0], cnt);
  }
}

// Lattice translation that brings each atom into the cell at the list build
// (image sums): W = floor(fractional coordinates), sorted order.
__global__ void k_wrap_shift(int n, const int32_t *__restrict__ perm, const double *__restrict__ xyz, CellD cell,
                             int32_t *__restrict__ wrap) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  double f0 = x * cell.A00 + y * cell.A10 + z * cell.A20, f1 = y * cell.A11 + z * cell.A21, f2 = z * cell.A22;
  wrap[3 * s] = (int)floor(f0);
  wrap[3 * s + 1] = (int)floor(f1);
  wrap[3 * s + 2] = (int)floor(f2);
}

// Full rows over atom images for cells thinner than twice the cutoff: every
// partner t (the atom itself included) at every lattice translation T with
// |x_t + T - x_s| <= cut, T in [-m, m] per axis (positions wrapped into the
// cell, so fractional differences lie in (-1, 1)).  Bonded and group codes
// apply only to the pair as given in the caller's coordinates, T = W_t - W_s.
// One thread per atom scans partners and translations in a fixed order.
__global__ void k_build_list_images(int n, const double4 *__restrict__ pos, const int32_t *__restrict__ perm,
                                    const int32_t *__restrict__ wrap, CellD cell, double cut2, int m0, int m1, int m2,
                                    const int32_t *__restrict__ group, const int32_t *__restrict__ moff,
                                    const int32_t *__restrict__ matom, const int32_t *__restrict__ mrel,
                                    uint32_t *__restrict__ nbr, int maxnb, int32_t *__restrict__ nnb,
                                    int32_t *__restrict__ flags) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double4 pi = pos[s];
  int wi0 = wrap[3 * s], wi1 = wrap[3 * s + 1], wi2 = wrap[3 * s + 2];
  int i = perm[s], gi = group[s];
  uint32_t *row = nbr + (size_t)s * maxnb;
  int cnt = 0;
  for (int t = 0; t < n; t++) {
    double4 pk = pos[t];
    double d0x = pk.x - pi.x, d0y = pk.y - pi.y, d0z = pk.z - pi.z;
    int o0 = wrap[3 * t] - wi0, o1 = wrap[3 * t + 1] - wi1, o2 = wrap[3 * t + 2] - wi2;
    int rel = -1;
    for (int tz = -m2; tz <= m2; tz++)
      for (int ty = -m1; ty <= m1; ty++)
        for (int tx = -m0; tx <= m0; tx++) {
          if (t == s && tx == 0 && ty == 0 && tz == 0) continue;
          double dx = d0x + tx * cell.Ai00 + ty * cell.Ai10 + tz * cell.Ai20;
          double dy = d0y + ty * cell.Ai11 + tz * cell.Ai21;
          double dz = d0z + tz * cell.Ai22;
          if (dx * dx + dy * dy + dz * dz > cut2) continue;
          int code = 0;
          if (tx == o0 && ty == o1 && tz == o2) {
            if (rel < 0) rel = relation(moff, matom, mrel, i, perm[t]);
            code = rel + (group[t] == gi ? 8 : 0);
          }
          uint32_t img = (uint32_t)((tx + NB_IMG_R) + NB_IMG_W * ((ty + NB_IMG_R) + NB_IMG_W * (tz + NB_IMG_R)));
          if (cnt < maxnb) row[cnt] = (uint32_t)t | (img << NB_IMG_SHIFT) | ((uint32_t)code << NB_CODE_SHIFT);
          cnt++;
        }
  }
  nnb[s] = cnt;
  if (cnt > maxnb) atomicMax(&flags[0], cnt);
}

}  // namespace

static int estimate_neighbors(const NbHandle *h, double r) {
  double vol = h->cell.aperiodic ? 1.0 : h->cell.vol;
  double rho = h->cell.aperiodic ? 0.12 : h->n / vol;
  double v = 4.0 / 3.0 * M_PI * r * r * r;
  double est = rho * v * 1.4 + 64.0;
  if (est > h->n) est = h->n;
  return est < 16.0 ? 16 : (int)est;
}

// recorded form of the list check: flags[3] (cleared by k_zero_outputs just
// before) is raised when an atom moved more than half the buffer; the caller
// reads it back with the energies
void nb_check_rebuild_async(NbHandle *h) {
  int n = h->n;
  double lim = 0.5 * h->prm.list_buffer;
  NB_LAUNCH(k_max_disp, dim3((n + 255) / 256), dim3(256), 0, n, h->d_xyz, h->d_xyz_list, lim * lim, h->d_flags);
}

bool nb_need_rebuild(NbHandle *h) {
  if (!h->list_valid) return true;
  int n = h->n;
  int32_t zero = 0;
  HIPCHECK(hipMemcpyAsync(h->d_flags + 3, &zero, sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
  double lim = 0.5 * h->prm.list_buffer;
  NB_LAUNCH(k_max_disp, dim3((n + 255) / 256), dim3(256), 0, n, h->d_xyz, h->d_xyz_list, lim * lim,
                     h->d_flags);
  int32_t flag = 0;
  HIPCHECK(hipMemcpyAsync(&flag, h->d_flags + 3, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return flag != 0;
}

void nb_build_lists(NbHandle *h) {
  int n = h->n;
  double cut = (h->prm.vdw_cutoff > h->prm.ewald_cutoff ? h->prm.vdw_cutoff : h->prm.ewald_cutoff) + h->prm.list_buffer;
  // cell grid
  int nc[3] = {1, 1, 1};
  if (!h->cell.aperiodic) {
    for (int d = 0; d < 3; d++) {
      double col = sqrt(h->cell.A[0][d] * h->cell.A[0][d] + h->cell.A[1][d] * h->cell.A[1][d] +
                        h->cell.A[2][d] * h->cell.A[2][d]);
      int q = (int)floor(1.0 / (col * cut));
      nc[d] = q < 1 ? 1 : (q > 128 ? 128 : q);
    }
  }
  int m = nc[0] * nc[1] * nc[2];
  if (m + 1 > h->ncell_cap) {
    if (h->d_cellcount) { hipFree(h->d_cellcount); hipFree(h->d_cellstart); }
    h->ncell_cap = m + 1;
    HIPCHECK(hipMalloc(&h->d_cellcount, sizeof(int32_t) * (m + 1)));
    HIPCHECK(hipMalloc(&h->d_cellstart, sizeof(int32_t) * (m + 1)));
  }
  h->ncell[0] = nc[0]; h->ncell[1] = nc[1]; h->ncell[2] = nc[2];
  h->ncell_tot = m;
  dim3 B(256), G((n + 255) / 256);
  HIPCHECK(hipMemsetAsync(h->d_cellcount, 0, sizeof(int32_t) * (m + 1), h->stream));
  NB_LAUNCH(k_cell_assign, G, B, 0, n, h->d_xyz, h->cell, nc[0], nc[1], nc[2], h->d_cellid,
                     h->d_cellcount);
  NB_LAUNCH(k_scan, dim3(1), dim3(1024), 0, m, h->d_cellcount, h->d_cellstart);
  HIPCHECK(hipMemsetAsync(h->d_cellcount, 0, sizeof(int32_t) * (m + 1), h->stream));
  // unsorted cell segments go to iperm (rebuilt just after), then are ranked into perm
  NB_LAUNCH(k_cell_scatter, G, B, 0, n, h->d_cellid, h->d_cellstart, h->d_cellcount, h->d_iperm);
  NB_LAUNCH(k_cell_rank, dim3(m), B, 0, m, h->d_cellstart, h->d_iperm, h->d_perm, h->d_xyz, h->cell, nc[0], nc[1],
            nc[2]);
  NB_LAUNCH(k_iperm, G, B, 0, n, h->d_perm, h->d_iperm);
  NB_LAUNCH(k_gather_static, G, B, 0, n, h->d_perm, h->d_polar0, h->d_thole0, h->d_ipd0,
                     h->d_group0, h->d_vclass0, h->d_soft0, h->d_mol0, h->d_polar, h->d_group, h->d_vclass, h->d_soft,
                     h->d_mol, h->d_mp);
  int mimg[3] = {0, 0, 0};
  if (h->cell.multi) {
    NB_LAUNCH(k_wrap_shift, G, B, 0, n, h->d_perm, h->d_xyz, h->cell, h->d_wrap);
    for (int d = 0; d < 3; d++) {
      const double(*A)[3] = h->cell.A;
      double w = 1.0 / sqrt(A[0][d] * A[0][d] + A[1][d] * A[1][d] + A[2][d] * A[2][d]);
      mimg[d] = (int)floor(cut / w) + 1;
      if (mimg[d] > NB_IMG_R) throw std::runtime_error("cell too thin for image sums at this cutoff");
    }
  }
  nb_gather_sorted(h);
  // neighbor rows, grown on overflow
  if (h->maxnb == 0) {
    h->maxnb = estimate_neighbors(h, cut);
    if (h->cell.multi) {
      // every partner may appear at several translations
      double est = h->n / h->cell.vol * 4.0 / 3.0 * M_PI * cut * cut * cut * 1.4 + 64.0;
      h->maxnb = (int)std::min(est, (double)n * (2 * mimg[0] + 1) * (2 * mimg[1] + 1) * (2 * mimg[2] + 1));
    }
  }
  for (int attempt = 0; attempt < 4; attempt++) {
    if (!h->d_nbr) HIPCHECK(hipMalloc(&h->d_nbr, sizeof(uint32_t) * (size_t)h->maxnb * n));
    int32_t zeros[4] = {0, 0, 0, 0};
    HIPCHECK(hipMemcpyAsync(h->d_flags, zeros, sizeof(zeros), hipMemcpyHostToDevice, h->stream));
    if (h->cell.multi)
      NB_LAUNCH(k_build_list_images, grid_for(n, 64), dim3(64), 0, n, h->d_pos4, h->d_perm, h->d_wrap, h->cell,
                cut * cut, mimg[0], mimg[1], mimg[2], h->d_group, h->d_mask_off, h->d_mask_atom, h->d_mask_rel,
                h->d_nbr, h->maxnb, h->d_nnb, h->d_flags);
    else
      NB_LAUNCH(k_build_list<16>, grid_for((long)n * 16, 256), B, 0, n, h->d_pos4, h->d_perm, h->d_cellstart, h->cell,
                nc[0], nc[1], nc[2], cut * cut, h->d_group, h->d_mask_off, h->d_mask_atom, h->d_mask_rel, h->d_nbr,
                h->maxnb, h->d_nnb, h->d_flags);
    int32_t f0 = 0;
    HIPCHECK(hipMemcpyAsync(&f0, h->d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (f0 <= h->maxnb) break;
    if (attempt == 3) throw std::runtime_error("neighbor list capacity could not be established");
    hipFree(h->d_nbr);
    h->d_nbr = nullptr;
    h->maxnb = f0 + f0 / 4 + 32;
  }
  // Ewald / preconditioner rows sized from the list row
  if (h->maxew == 0) {
    h->maxew = estimate_neighbors(h, h->prm.ewald_cutoff);
    if (h->maxew > h->maxnb) h->maxew = h->maxnb;
    h->maxpc = estimate_neighbors(h, h->prm.precond_cutoff);
    if (h->maxpc > h->maxew) h->maxpc = h->maxew;
    HIPCHECK(hipMalloc(&h->d_ew, sizeof(uint32_t) * (size_t)h->maxew * n));
    HIPCHECK(hipMalloc(&h->d_pc, sizeof(uint32_t) * (size_t)h->maxpc * n));
  }
  HIPCHECK(hipMemcpyAsync(h->d_xyz_list, h->d_xyz, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, h->stream));
  h->list_valid = true;
}
