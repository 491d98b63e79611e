// cache.hip -- frustum brick cache: blockwise fill (CacheRenderer::
// initCacheBlockwise, src/cache_renderer.cu:18-569, 931-1051) and the
// reprojecting cache renderer (CachePerformanceRenderer::render,
// src/cache_performance_renderer.cu:22-313, 670-935, 1059-1263).
//
// Storage: a sparse brick pool of 8^3-voxel bricks (32-byte voxel records:
// 8 fp16 latents, f32 density, f32 isset) addressed through an int32
// indirection table over the frustum's brick grid (-1 = unset).  The
// reference's dense 1024x1024xD 3D arrays and texture objects are replaced
// by this table + in-kernel point fetches; trilinear filtering is done in
// software exactly as its live path does (cache_performance_renderer.cu:678-787).
//
// Fill: one wave64 per 8x8 pixel tile (the reference's 64-thread cg tile).
// Per sub-iteration every alive tile processes its next non-empty 8-step
// slab.  Brick slots are handed out by a scan over the alive tiles in tile
// order, which makes the indirection table independent of scheduling (the
// reference's atomicAdd at cache_renderer.cu:420 is run-dependent).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "march.h"
#include "mlp_mfma.h"
#include "render_common.h"

namespace fvc {

struct FillRay {
  float ox, oy, oz, dx, dy, dz, far_, T;
  int alive, cnt, off, pad;
};
struct FillTile {
  int slab, alive, n_bricks, processed;
  int visible, slot, overflow, pad;
};

// ------------------------------------------------------------------ fill
// initPayloads_cache_init_kernel (cache_renderer.cu:18-81)
__global__ __launch_bounds__(64) void k_fill_init(fvc_raymarch_info rm, V3 aabb_from, V3 aabb_to, int bx,
                                                  FillRay* __restrict__ rays, FillTile* __restrict__ tiles,
                                                  uint8_t* __restrict__ alive8) {
  const int t = blockIdx.x, lane = threadIdx.x;
  const int px = (t % bx) * BRICK + (lane & 7), py = (t / bx) * BRICK + (lane >> 3);
  const fvc_stepsize_info& s = rm.stepsize_info;
  const Ray r = gen_ray(rm.cam_info, (float)px + 0.5f, (float)py + 0.5f);
  float a0, a1;
  const bool hit = ray_aabb(r, aabb_from, aabb_to, a0, a1);
  FillRay fr;
  fr.ox = r.o.x; fr.oy = r.o.y; fr.oz = r.o.z;
  fr.dx = r.d.x; fr.dy = r.d.y; fr.dz = r.d.z;
  fr.far_ = hit ? fminf_(a1, s.far_) : s.far_;
  fr.T = 1.0f;
  fr.alive = hit;
  fr.cnt = 0; fr.off = 0; fr.pad = 0;
  rays[(int64_t)t * 64 + lane] = fr;
  int start = (int)(step_from_t(s, hit ? a0 : s.near_) - s.near_i);
  start = start > 0 ? start : 0;
  const int mn = wave_min(start);
  const int n_alive = wave_sum(hit ? 1 : 0);
  if (lane == 0) {
    FillTile ft{};
    ft.slab = mn / BRICK;
    ft.alive = n_alive > 0;
    tiles[t] = ft;
    alive8[t] = (uint8_t)ft.alive;
  }
}

// Fill counters (Renderer::f_counters): CNT_* slots, then per sub-iteration
// {alive tiles, samples} from FILL_LOG on.
static const int FILL_LOG = 64;
static const int FILL_MAX_ITERS = 1000;  // CacheRenderer::MAX_SUBITER
static const int FILL_COUNTER_WORDS = FILL_LOG + 2 * (FILL_MAX_ITERS + 1);

__global__ void k_set_ints(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// alive tiles in tile order (deterministic compaction); one block, each
// thread takes a run of consecutive tiles (flags in a bit mask), one block
// prefix count over the runs.  The flags are the tiles' alive bytes (alive8,
// kept beside the tile records), read coalesced.
static const int SCAN_RUN = 32;  // tiles per thread: 32768 tiles per pass of 1024 threads
__device__ __forceinline__ int block_excl_sum(int v, int& total) {
  __shared__ int s_w[16];
  const int wave = (int)(threadIdx.x >> 6), n_waves = (int)(blockDim.x >> 6);
  int wt;
  const int pre = wave_excl_scan(v, wt);
  if (lane_id() == 0) s_w[wave] = wt;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < n_waves; w++) {
    const int c = s_w[w];
    off += w < wave ? c : 0;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return off + pre;
}

__global__ __launch_bounds__(1024) void k_fill_list(int n_tiles, const uint8_t* __restrict__ alive8,
                                                    int* __restrict__ list, int* __restrict__ cnt, int it) {
  __shared__ uint8_t flags[1024 * SCAN_RUN];
  int base = 0;
  for (int c0 = 0; c0 < n_tiles; c0 += 1024 * SCAN_RUN) {
    // coalesced loads (independent, all in flight), then runs from LDS
#pragma unroll
    for (int k = 0; k < SCAN_RUN; k++) {
      const int i = c0 + k * 1024 + (int)threadIdx.x;
      flags[k * 1024 + threadIdx.x] = i < n_tiles ? alive8[i] : (uint8_t)0;
    }
    __syncthreads();
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < SCAN_RUN; k++) bits |= (uint32_t)flags[threadIdx.x * SCAN_RUN + k] << k;
    const int b = c0 + (int)threadIdx.x * SCAN_RUN;
    int tot;
    int o = base + block_excl_sum(__popc(bits), tot);
    while (bits) {
      const int k = __ffs(bits) - 1;
      bits &= bits - 1;
      list[o++] = b + k;
    }
    base += tot;
  }
  if (threadIdx.x == 0) {
    cnt[CNT_TILES] = base;
    cnt[CNT_SAMPLES] = 0;
    if (it <= FILL_MAX_ITERS) cnt[FILL_LOG + 2 * it] = base;
  }
}

// sample_cache_init_first_kernel + sample_cache_init_second_kernel
// (cache_renderer.cu:83-287): skip empty slabs, emit the occupied samples of
// the first non-empty one as a compacted stream.  One wave per tile of the
// alive list (grid-stride).  After an empty slab the wave jumps over the
// slabs that every live lane's empty box covers: a lane's first bin that is
// neither in its skip interval nor past its far bound lies at or after
// step_from_t(min(hi, far)) - near_i - 2 (two bins of margin for rounding),
// so no slab before the wave minimum of that bound (in slabs) holds a
// sample or ends a ray, and the tile ends at the same slab as a walk.
__global__ __launch_bounds__(64) void k_fill_sample(const int* __restrict__ list, const int* __restrict__ cnt,
                                                    fvc_stepsize_info st, OccParams occ, ContractParams con,
                                                    const uint64_t* __restrict__ bf,
                                                    const uint32_t* __restrict__ coarse_g, int bz,
                                                    FillRay* __restrict__ rays, FillTile* __restrict__ tiles,
                                                    float* __restrict__ f_pos, float* __restrict__ f_t0,
                                                    int* __restrict__ f_meta, int* __restrict__ counter,
                                                    float* __restrict__ f_dir, uint8_t* __restrict__ alive8) {
  const int n_list = cnt[CNT_TILES];
  if ((int)blockIdx.x >= n_list) return;
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, occ.nlvl, coarse);
  const int lane = threadIdx.x;
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    const int t = list[li];
    FillRay r = rays[(int64_t)t * 64 + lane];
    FillTile tile = tiles[t];
    const V3 o = {r.ox, r.oy, r.oz}, d = {r.dx, r.dy, r.dz};
    const RaySlab rslab = ray_slab(o, d);
    Skip sk = no_skip();
    int slab = tile.slab;
    unsigned mask = 0;
    int total = 0;
    while (true) {
      if (wave_sum(r.alive) <= 0) break;
      if (slab >= bz) {  // beyond the frustum: the column ends
        r.alive = 0;
        break;
      }
      float t1 = t_from_step(st, (float)(slab * BRICK) + st.near_i);
      float tms[BRICK];
#pragma unroll
      for (int j = 0; j < BRICK; j++) {
        const float t0 = t1;
        t1 = t0 + step_dt(st, t0);
        tms[j] = (t0 + t1) * 0.5f;
      }
      // bins inside the skip interval of an earlier empty bin need no query
      mask = 0;
#pragma unroll
      for (int j = 0; j < BRICK; j++) {
        const float tm = tms[j];
        if (tm > r.far_) { r.alive = 0; continue; }
        if (in_skip(sk, tm)) continue;
        if (occ_query(o + tm * d, occ, bf, coarse, rslab, sk)) mask |= 1u << j;
      }
      total = wave_sum(__popc(mask));
      if (total > 0) break;
      slab++;
      if (wave_sum(r.alive) > 0) {
        int cand = 0x7fffffff;
        if (r.alive) {
          const float hi = fminf_(sk.hi, r.far_);
          const float sl = fminf_((step_from_t(st, hi) - st.near_i - 2.0f) * 0.125f, 1.0e9f);
          cand = sk.hi >= sk.lo ? (int)floorf(sl) : slab;
        }
        const int nxt = wave_min(cand);
        if (nxt > slab) slab = nxt < bz ? nxt : bz;
      }
    }
    const bool tile_alive = wave_sum(r.alive) > 0;
    const bool processed = total > 0 && tile_alive;
    r.cnt = 0;
    if (processed) {
      const int n = __popc(mask);
      const int off = wave_reserve(n, counter);
      float t1 = t_from_step(st, (float)(slab * BRICK) + st.near_i);
      int k = 0;
      for (int j = 0; j < BRICK; j++) {
        const float t0 = t1;
        const float dt = step_dt(st, t0);
        t1 = t0 + dt;
        if (mask & (1u << j)) {
          const float tm = (t0 + t1) * 0.5f;
          const V3 p = contract(o + tm * d, con);
          const int64_t q = (int64_t)off + k;
          f_pos[3 * q] = p.x; f_pos[3 * q + 1] = p.y; f_pos[3 * q + 2] = p.z;
          f_t0[q] = t0;
          f_meta[q] = j;
          if (f_dir) {  // split models: the fill ray is the initial view direction (cache_renderer.cu:899)
            const V3 d01 = unit_to_01(d);
            f_dir[3 * q] = d01.x; f_dir[3 * q + 1] = d01.y; f_dir[3 * q + 2] = d01.z;
          }
          k++;
        }
      }
      r.cnt = n;
      r.off = off;
    }
    rays[(int64_t)t * 64 + lane] = r;
    if (lane == 0) {
      tile.slab = slab;
      tile.alive = tile_alive;
      tile.processed = processed;
      tile.visible = 0;
      tiles[t] = tile;
      alive8[t] = (uint8_t)tile_alive;
    }
  }
}

// half in row r of a sample's 16-half network output
__device__ __forceinline__ float out16_row(const uint2* out16, int64_t q, int r) {
  const uint2 u = out16[q * 4 + (r >> 2)];
  const uint32_t w = (r & 2) ? u.y : u.x;
  return h2f((uint16_t)((r & 1) ? w >> 16 : w & 0xffffu));
}

// first loop of accumulateNew_cache_init_second_kernel (cache_renderer.cu:358-405):
// transmittance over the slab's samples and the count of visible samples.
// sig_row: row of log sigma in the network output (0; split models 8);
// lsig (plain models): that row as a compact half array, read instead.
// Logs the sub-iteration's sample count.
__global__ __launch_bounds__(64) void k_fill_accum(const int* __restrict__ list, int* __restrict__ cnt, int it,
                                                   fvc_stepsize_info st, float alpha_thre,
                                                   const uint2* __restrict__ out16, const float* __restrict__ f_t0,
                                                   FillRay* __restrict__ rays, FillTile* __restrict__ tiles,
                                                   int sig_row, uint8_t* __restrict__ alloc8,
                                                   const uint16_t* __restrict__ lsig) {
  const int n_list = cnt[CNT_TILES];
  if (blockIdx.x == 0 && threadIdx.x == 0 && it <= FILL_MAX_ITERS) cnt[FILL_LOG + 2 * it + 1] = cnt[CNT_SAMPLES];
  const int lane = threadIdx.x;
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    const int t = list[li];
    if (!tiles[t].processed) {
      if (lane == 0) alloc8[li] = 0;
      continue;
    }
    FillRay r = rays[(int64_t)t * 64 + lane];
    int vis = 0;
    // a lane has at most 8 samples (one slab): all loaded before the
    // sequential transmittance product
    float lsv[BRICK], t0v[BRICK];
#pragma unroll
    for (int k = 0; k < BRICK; k++) {
      // loads from a valid index (0 past the lane's samples)
      const int64_t q = k < r.cnt ? (int64_t)r.off + k : 0;
      lsv[k] = lsig ? h2f(lsig[q]) : out16_row(out16, q, sig_row);
      t0v[k] = f_t0[q];
    }
#pragma unroll
    for (int k = 0; k < BRICK; k++) {
      if (k >= r.cnt) break;
      const float ls = lsv[k];
      const float t0 = t0v[k];
      const float dt = step_dt(st, t0);
      const float density = expf(ls);
      float alpha = 1.0f - expf(-density * dt);
      alpha = fminf_(fmaxf_(alpha, 0.0f), 1.0f);
      if (alpha > alpha_thre) {
        r.T = r.T * (1.0f - alpha);
        vis++;
      }
      if (r.T < 1e-4f) r.alive = 0;
    }
    rays[(int64_t)t * 64 + lane].T = r.T;
    rays[(int64_t)t * 64 + lane].alive = r.alive;
    const int v = wave_sum(vis);
    if (lane == 0) {
      tiles[t].visible = v;
      alloc8[li] = (uint8_t)(v > 0);  // the tile takes a brick this sub-iteration
    }
  }
}

// brick slot allocation in list (= tile) order; one block, runs of
// consecutive list entries per thread, one block prefix count per pass
__global__ __launch_bounds__(1024) void k_fill_alloc(const int* __restrict__ list, int* __restrict__ cnt,
                                                     int capacity, FillTile* __restrict__ tiles,
                                                     const uint8_t* __restrict__ alloc8) {
  __shared__ uint8_t flags[1024 * SCAN_RUN];
  const int n = cnt[CNT_TILES];
  int base = cnt[CNT_SLOTS];
  bool any_overflow = false;
  for (int c0 = 0; c0 < n; c0 += 1024 * SCAN_RUN) {
#pragma unroll
    for (int k = 0; k < SCAN_RUN; k++) {
      const int i = c0 + k * 1024 + (int)threadIdx.x;
      // processed with visible samples (k_fill_accum's byte per list entry)
      flags[k * 1024 + threadIdx.x] = i < n ? alloc8[i] : (uint8_t)0;
    }
    __syncthreads();
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < SCAN_RUN; k++) bits |= (uint32_t)flags[threadIdx.x * SCAN_RUN + k] << k;
    const int b = c0 + (int)threadIdx.x * SCAN_RUN;
    int tot;
    int o = base + block_excl_sum(__popc(bits), tot);
    while (bits) {
      const int k = __ffs(bits) - 1;
      bits &= bits - 1;
      const int t = list[b + k];
      tiles[t].slot = o;
      tiles[t].overflow = o >= capacity;
      any_overflow |= o >= capacity;
      o++;
    }
    base += tot;
  }
  if (any_overflow) cnt[CNT_OVERFLOW] = 1;
  __syncthreads();
  if (threadIdx.x == 0) cnt[CNT_SLOTS] = base;
}

// brick write + slab advance (accumulateNew_cache_init_second_kernel,
// cache_renderer.cu:405-528).  Each lane owns one (x,y) column of the brick
// and produces its 8 voxels (zeros where it has no sample); the brick is
// assembled in LDS and stored as contiguous 16-byte chunks.
// With FP8 latents (pool8 / scales) the brick's scale is the absmax of the
// latents written into it over 448 (1 when all are zero); each latent is
// stored as the E4M3 code (round to nearest even) of l / scale.
// Latents come from rows latent_offset.., log sigma from row sig_row; di
// (DensityIntermediates): latent 0 is replaced by the density as half
// (cache_renderer.cu:481-486, 501).
__global__ __launch_bounds__(64) void k_fill_write(const int* __restrict__ list, const int* __restrict__ cnt, int bx,
                                                   int by, int latent_offset, const uint16_t* __restrict__ out16,
                                                   const int* __restrict__ f_meta, const FillRay* __restrict__ rays,
                                                   FillTile* __restrict__ tiles, int* __restrict__ table,
                                                   Voxel* __restrict__ pool, Voxel8* __restrict__ pool8,
                                                   float* __restrict__ scales, int sig_row, int di,
                                                   uint8_t* __restrict__ alive8) {
  __shared__ uint4 brick_lds[BRICK_VOXELS * sizeof(Voxel) / 16];
  const int n_list = cnt[CNT_TILES];
  const int lane = threadIdx.x;
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    const int t = list[li];
    FillTile tile = tiles[t];
    if (!tile.processed) continue;
    const FillRay r = rays[(int64_t)t * 64 + lane];
    bool overflow = false;
    if (tile.visible > 0) {
      overflow = tile.overflow != 0;
      if (lane == 0) table[((int64_t)tile.slab * by + t / bx) * bx + t % bx] = overflow ? -1 : tile.slot;
      // the lane's brick column (x, y) = (lane & 7, lane >> 3) is 8
      // consecutive records starting at brick_voxel(x, y, 0)
      const int col = brick_voxel(lane & 7, lane >> 3, 0);
      if (!overflow && pool8) {
        // latent i of a sample's output (fp16 value)
        auto lat = [&](const uint16_t* o16, int i) {
          return (di && i == 0) ? (float)(_Float16)expf(h2f(o16[sig_row])) : h2f(o16[latent_offset + i]);
        };
        int amax_bits = 0;  // |latent| as float bits (non-negative floats order as integers)
        for (int k = 0; k < r.cnt; k++) {
          const uint16_t* o16 = out16 + ((int64_t)r.off + k) * 16;
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const int b = __builtin_bit_cast(int, fabsf(lat(o16, i)));
            amax_bits = b > amax_bits ? b : amax_bits;
          }
        }
        amax_bits = -wave_min(-amax_bits);
        const float amax = __builtin_bit_cast(float, amax_bits);
        const float scale = amax > 0.0f ? amax / 448.0f : 1.0f;
        if (lane == 0) scales[tile.slot] = scale;
        Voxel8* lds8 = reinterpret_cast<Voxel8*>(brick_lds);
        int k = 0;
        for (int z = 0; z < BRICK; z++) {
          Voxel8 v;
          const bool has = k < r.cnt && f_meta[r.off + k] == z;
          if (has) {
            const int64_t q = (int64_t)r.off + k;
            const uint16_t* o16 = out16 + q * 16;
            float qv[8];
#pragma unroll
            for (int i = 0; i < 8; i++) qv[i] = fminf(fmaxf(lat(o16, i) / scale, -448.0f), 448.0f);
            int lo = __builtin_amdgcn_cvt_pk_fp8_f32(qv[0], qv[1], 0, false);
            lo = __builtin_amdgcn_cvt_pk_fp8_f32(qv[2], qv[3], lo, true);
            int hi = __builtin_amdgcn_cvt_pk_fp8_f32(qv[4], qv[5], 0, false);
            hi = __builtin_amdgcn_cvt_pk_fp8_f32(qv[6], qv[7], hi, true);
            v.lat[0] = (uint32_t)lo;
            v.lat[1] = (uint32_t)hi;
            v.density = expf(h2f(o16[sig_row]));
            v.isset = 1.0f;
            k++;
          } else {
            v.lat[0] = v.lat[1] = 0;
            v.density = 0.0f;
            v.isset = 0.0f;
          }
          lds8[col + z] = v;
        }
        wave_sync();
        uint4* dst = reinterpret_cast<uint4*>(pool8 + (int64_t)tile.slot * BRICK_VOXELS);
        for (int i = lane; i < BRICK_VOXELS * (int)sizeof(Voxel8) / 16; i += 64) dst[i] = brick_lds[i];
        wave_sync();
        tile.n_bricks++;
      } else if (!overflow) {
        // the lane's column in LDS: zeros, then a record at the z of each
        // sample (the lane's at most 8 samples loaded together), then each
        // record's znext from the one above it
        Voxel* lds = reinterpret_cast<Voxel*>(brick_lds);
        const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int z = 0; z < BRICK; z++) {
          reinterpret_cast<uint4*>(lds + col + z)[0] = zero;
          reinterpret_cast<uint4*>(lds + col + z)[1] = zero;
        }
        int meta[BRICK];
        uint4 row[BRICK][2];
#pragma unroll
        for (int k = 0; k < BRICK; k++) {
          // loads from a valid index (0 past the lane's samples), selected after
          const int64_t q = k < r.cnt ? (int64_t)r.off + k : 0;
          const int mt = f_meta[q];
          const uint4 r0 = reinterpret_cast<const uint4*>(out16 + q * 16)[0];
          const uint4 r1 = reinterpret_cast<const uint4*>(out16 + q * 16)[1];
          meta[k] = k < r.cnt ? mt : -1;
          row[k][0] = r0;
          row[k][1] = r1;
        }
#pragma unroll
        for (int k = 0; k < BRICK; k++) {
          if (meta[k] < 0) break;
          // halves 0..8 of the output row (latents start at row 0 or 1, log
          // sigma is row 0 or 8), taken at fixed positions
          const uint32_t w[5] = {row[k][0].x, row[k][0].y, row[k][0].z, row[k][0].w, row[k][1].x};
          uint16_t e[10];
#pragma unroll
          for (int m = 0; m < 10; m++) e[m] = (uint16_t)(w[m >> 1] >> (16 * (m & 1)));
          Voxel v;
#pragma unroll
          for (int i = 0; i < 8; i++) v.lat[i] = latent_offset ? e[i + 1] : e[i];
          v.density = expf(h2f(sig_row ? e[8] : e[0]));
          if (di) v.lat[0] = __builtin_bit_cast(uint16_t, (_Float16)v.density);
          v.isset = 1.0f;
          v.znext[0] = v.znext[1] = 0.0f;
          lds[col + meta[k]] = v;
        }
#pragma unroll
        for (int z = 0; z < BRICK - 1; z++) {
          lds[col + z].znext[0] = lds[col + z + 1].density;
          lds[col + z].znext[1] = lds[col + z + 1].isset;
        }
        wave_sync();
        uint4* dst = reinterpret_cast<uint4*>(pool + (int64_t)tile.slot * BRICK_VOXELS);
        for (int i = lane; i < BRICK_VOXELS * (int)sizeof(Voxel) / 16; i += 64) dst[i] = brick_lds[i];
        wave_sync();
        tile.n_bricks++;
      }
    }
    const int n_alive = wave_sum(r.alive);
    if (lane == 0) {
      if (!overflow) tile.slab++;
      tile.alive = !overflow && n_alive > 0;
      tiles[t] = tile;
      alive8[t] = (uint8_t)tile.alive;
    }
  }
}

// writeImageBuffer_cache_init_kernel2 (cache_renderer.cu:530-569): known range
__global__ __launch_bounds__(64) void k_fill_known(int W, int H, int bx, fvc_stepsize_info st,
                                                   const FillTile* __restrict__ tiles, float* __restrict__ known) {
  const int t = blockIdx.x, lane = threadIdx.x;
  const int px = (t % bx) * BRICK + (lane & 7), py = (t / bx) * BRICK + (lane >> 3);
  if (px < W && py < H) known[(int64_t)py * W + px] = t_from_step(st, (float)(tiles[t].slab * BRICK) + st.near_i);
}

// ------------------------------------------------------------------ geometry
// Frustum depth in steps: the reference's max_step_in_scene
// (raymarch_common.h:298-303, cache_renderer.cu:841-843), ceil(step_from_t(
// |AABB diagonal| + near)) padded to 4.  A camera inside the AABB sees no
// point farther than the diagonal, so this is the reference's depth exactly.
// A fill camera outside the AABB can see points up to its farthest AABB
// corner; the reference fill has no depth guard there (its slab walk,
// cache_renderer.cu:130-178, would address bricks past the frustum), so the
// diagonal is replaced by that corner's distance (at most far) whenever it is
// longer, and every brick stays inside the frustum.
static int frustum_depth(const fvc_raymarch_info& rm, const fvc_scene_info& sc) {
  const fvc_stepsize_info& st = rm.stepsize_info;
  const V3 from = cvt(sc.aabb_from), to = cvt(sc.aabb_to);
  const V3 dg = to - from;
  const float diag = std::sqrt(dot3(dg, dg));
  const V3 o = translation(rm.cam_info.cam2world);
  float tf = 0.0f;
  for (int c = 0; c < 8; c++) {
    const V3 corner = {(c & 1) ? to.x : from.x, (c & 2) ? to.y : from.y, (c & 4) ? to.z : from.z};
    const V3 v = corner - o;
    tf = std::max(tf, std::sqrt(dot3(v, v)));
  }
  tf = std::min(tf, st.far_);
  const float reach = tf > diag ? tf : diag;
  const int steps = (int)std::ceil(step_from_t(st, reach + st.near_));
  return (steps + 3) / 4 * 4;
}

// CacheRenderer::resizeCache estimate (cache_renderer.cu:832-846 + cache.cu:143-146)
static int default_capacity(int64_t n_rays) {
  const int64_t est_samples = n_rays * 32;
  const int64_t per_brick = (int64_t)(0.5f * BRICK * BRICK * BRICK);
  const int64_t est_bricks = est_samples / per_brick;
  const int64_t per_layer = 128 * 128;
  return (int)std::min<int64_t>(MAX_BRICKS, (est_bricks + per_layer - 1) / per_layer * per_layer);
}

// CachePerformanceRenderer's constructor check (cache_performance_renderer.cu:
// 964-967): the cache holds 8 latents (2 half4 arrays)
static void check_cache_latents(const Model& m) {
  if (m.net.n_latents > 8)
    throw Error(FVC_E_UNSUPPORTED,
                "Latent width of the model is larger than the (hard-coded) cache size of 8 latents");
}

// The fill keeps its own counters, pinned mirror and scratch, so it can run
// on the fill stream from another thread while frames render from the
// other cache (double-buffered renderers).
//
// Sub-iterations are planned on the device: every kernel reads the alive
// tile count and the sample count that the previous kernels left in the
// fill counters, and idles when nothing is left, so the host issues bursts
// of sub-iterations and reads the per-iteration log {alive tiles, samples}
// once per burst (the reference reads two counts per sub-iteration,
// cache_renderer.cu:986-991, 1003-1006).  The first burst is as long as the
// previous fill's loop, plus one; with graph mode it is replayed as a graph.
void cache_fill(Renderer& r, const fvc_raymarch_info& rm, fvc_init_stats* stats, hipStream_t s, Profiler& P) {
  const Model& m = *r.model;
  check_cache_latents(m);
  CacheStore& C = r.caches[r.fill_idx];
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  r.f_counters.ensure(FILL_COUNTER_WORDS * sizeof(int));
  if (!r.f_pinned)
    FVC_HIP(hipHostMalloc((void**)&r.f_pinned, FILL_COUNTER_WORDS * sizeof(int), hipHostMallocDefault));

  // resize (Cache::resize, cache.cu:44-84), sparse pool sized by the estimate
  C.rm = rm;
  C.fx = W; C.fy = H; C.fz = frustum_depth(rm, m.scene);
  C.bx = (W + BRICK - 1) / BRICK; C.by = (H + BRICK - 1) / BRICK; C.bz = (C.fz + BRICK - 1) / BRICK;
  C.capacity = std::min(MAX_BRICKS, r.capacity_override > 0 ? r.capacity_override : default_capacity((int64_t)W * H));
  const int64_t n_table = (int64_t)C.bx * C.by * C.bz;
  C.table.ensure(n_table * sizeof(int32_t));
  C.known.ensure((int64_t)W * H * sizeof(float));
  C.lat8 = r.latent_format == FVC_LATENT_FP8;
  C.pool.ensure(std::max<int64_t>(1, (int64_t)C.capacity) * BRICK_VOXELS *
                (C.lat8 ? sizeof(Voxel8) : sizeof(Voxel)));
  C.scales.ensure(std::max<int64_t>(1, (int64_t)C.capacity) * sizeof(float));
  // a double-buffered renderer's other cache, while it holds nothing, gets
  // buffers of this cache's sizes now: the first background fill into it
  // (a view of about this size) then makes no hipMalloc, which would stall
  // the frames rendering meanwhile
  if (r.double_buffered) {
    CacheStore& O = r.caches[r.fill_idx ^ 1];
    if (&O != &C && !O.initialized) {
      O.table.ensure(n_table * sizeof(int32_t));
      O.known.ensure((int64_t)W * H * sizeof(float));
      O.pool.ensure(std::max<int64_t>(1, (int64_t)C.capacity) * BRICK_VOXELS *
                    (C.lat8 ? sizeof(Voxel8) : sizeof(Voxel)));
      O.scales.ensure(std::max<int64_t>(1, (int64_t)C.capacity) * sizeof(float));
    }
  }

  const int n_tiles = C.bx * C.by;
  const int64_t f_cap = (int64_t)n_tiles * 512;
  r.f_rays.ensure((int64_t)n_tiles * 64 * sizeof(FillRay));
  r.f_tiles.ensure((int64_t)n_tiles * sizeof(FillTile));
  r.f_list.ensure((int64_t)n_tiles * sizeof(int));
  r.f_alive8.ensure(n_tiles);
  r.f_alloc8.ensure(n_tiles);
  r.f_pos.ensure(f_cap * 3 * sizeof(float));
  r.f_t0.ensure(f_cap * sizeof(float));
  r.f_meta.ensure(f_cap * sizeof(int));
  r.f_out16.ensure(f_cap * 16 * sizeof(uint16_t));
  r.f_lsig.ensure(f_cap * sizeof(uint16_t));
  const bool split = m.net.split_head != 0;
  if (split) r.f_dir.ensure(f_cap * 3 * sizeof(float));
  // rows of the network output: log sigma and the first stored latent
  const int sig_row = split ? 8 : 0, lat_row = split ? 0 : m.net.latent_offset;
  const int di = m.net.first_latent_is_density ? 1 : 0;
  int* cnt = r.f_counters.as<int>();
  // one block (wave) per tile of the alive list; blocks past the list's
  // length return at once (grid-stride beyond 65536 tiles)
  const unsigned grid_t = (unsigned)std::min(n_tiles, 65536);

  // clears are kernels, not memsets: the fill may be recorded into a graph
  auto setup = [&] {
    hipLaunchKernelGGL(k_set_ints, dim3((unsigned)((n_table + 255) / 256)), dim3(256), 0, s, C.table.as<int>(),
                       n_table, -1);
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(FILL_LOG), 0, s, cnt, (int64_t)FILL_LOG, 0);
    hipLaunchKernelGGL(k_fill_init, dim3(n_tiles), dim3(64), 0, s, rm, cvt(m.scene.aabb_from),
                       cvt(m.scene.aabb_to), C.bx, r.f_rays.as<FillRay>(), r.f_tiles.as<FillTile>(),
                       r.f_alive8.as<uint8_t>());
    FVC_HIP(hipGetLastError());
  };
  struct Handles { int list, sample, net, accum, alloc, write; };
  std::vector<Handles> hs;
  auto iteration = [&](int it) {
    Handles h;
    hipEvent_t ev = P.begin(s);
    hipLaunchKernelGGL(k_fill_list, dim3(1), dim3(1024), 0, s, n_tiles, r.f_alive8.as<uint8_t>(),
                       r.f_list.as<int>(), cnt, it);
    FVC_HIP(hipGetLastError());
    h.list = P.end(ev, s, "fill_list");
    ev = P.begin(s);
    hipLaunchKernelGGL(k_fill_sample, dim3(grid_t), dim3(64), 0, s, r.f_list.as<int>(), cnt, rm.stepsize_info, m.occ,
                       m.con, m.bitfield.as<uint64_t>(), m.coarse.as<uint32_t>(), C.bz, r.f_rays.as<FillRay>(),
                       r.f_tiles.as<FillTile>(), r.f_pos.as<float>(), r.f_t0.as<float>(), r.f_meta.as<int>(),
                       cnt + CNT_SAMPLES, split ? r.f_dir.as<float>() : nullptr, r.f_alive8.as<uint8_t>());
    FVC_HIP(hipGetLastError());
    h.sample = P.end(ev, s, "fill_sample");
    ev = P.begin(s);
    if (split)
      launch_network_latents_vd(m, r.f_pos.as<float>(), r.f_dir.as<float>(), f_cap, r.f_out16.as<uint16_t>(), s,
                                r.mlp_rows, cnt + CNT_SAMPLES);
    else
      launch_network_latents(m, r.f_pos.as<float>(), f_cap, r.f_out16.as<uint16_t>(), s, r.mlp_rows,
                             cnt + CNT_SAMPLES, r.f_lsig.as<uint16_t>());
    h.net = P.end(ev, s, "network_latents");
    ev = P.begin(s);
    hipLaunchKernelGGL(k_fill_accum, dim3(grid_t), dim3(64), 0, s, r.f_list.as<int>(), cnt, it, rm.stepsize_info,
                       m.scene.alpha_thre, r.f_out16.as<uint2>(), r.f_t0.as<float>(), r.f_rays.as<FillRay>(),
                       r.f_tiles.as<FillTile>(), sig_row, r.f_alloc8.as<uint8_t>(),
                       split ? nullptr : r.f_lsig.as<uint16_t>());
    FVC_HIP(hipGetLastError());
    h.accum = P.end(ev, s, "fill_accum");
    ev = P.begin(s);
    hipLaunchKernelGGL(k_fill_alloc, dim3(1), dim3(1024), 0, s, r.f_list.as<int>(), cnt, C.capacity,
                       r.f_tiles.as<FillTile>(), r.f_alloc8.as<uint8_t>());
    FVC_HIP(hipGetLastError());
    h.alloc = P.end(ev, s, "fill_alloc");
    ev = P.begin(s);
    hipLaunchKernelGGL(k_fill_write, dim3(grid_t), dim3(64), 0, s, r.f_list.as<int>(), cnt, C.bx, C.by, lat_row,
                       r.f_out16.as<uint16_t>(), r.f_meta.as<int>(), r.f_rays.as<FillRay>(),
                       r.f_tiles.as<FillTile>(), C.table.as<int>(), C.lat8 ? nullptr : C.pool.as<Voxel>(),
                       C.lat8 ? C.pool.as<Voxel8>() : nullptr, C.scales.as<float>(), sig_row, di,
                       r.f_alive8.as<uint8_t>());
    FVC_HIP(hipGetLastError());
    h.write = P.end(ev, s, "fill_write");
    hs.push_back(h);
  };
  // the log of sub-iterations [a, b): true when the loop has ended
  auto read_log = [&](int a, int b) {
    FVC_HIP(hipMemcpyAsync(r.f_pinned + FILL_LOG + 2 * a, cnt + FILL_LOG + 2 * a, 2 * (b - a) * sizeof(int),
                           hipMemcpyDeviceToHost, s));
    FVC_HIP(hipStreamSynchronize(s));
    for (int i = a; i < b; i++)
      if (r.f_pinned[FILL_LOG + 2 * i] == 0 || r.f_pinned[FILL_LOG + 2 * i + 1] == 0) return true;
    return false;
  };
  int it = 0;
  bool done = false;
  const int first = std::min(FILL_MAX_ITERS, std::max(2, r.last_fill_iters + 1));
  const bool graph = r.use_graph && r.last_fill_iters > 0 && !P.enabled;
  if (graph) {
    graph_run(r, 2, s, [&] {
      setup();
      for (int i = 0; i < first; i++) iteration(i);
    });
  } else {
    setup();
    for (; it < first; it++) iteration(it);
  }
  it = first;
  done = read_log(0, it);
  while (!done && it < FILL_MAX_ITERS) {
    const int a = it;
    for (int b = 0; b < 4 && it < FILL_MAX_ITERS; b++, it++) iteration(it);
    done = read_log(a, it);
  }
  hipEvent_t ev = P.begin(s);
  hipLaunchKernelGGL(k_fill_known, dim3(n_tiles), dim3(64), 0, s, W, H, C.bx, rm.stepsize_info,
                     r.f_tiles.as<FillTile>(), C.known.as<float>());
  FVC_HIP(hipGetLastError());
  P.end(ev, s, "fill_known", (double)W * H, (double)W * H * 4 + n_tiles * 32.0, 0);
  FVC_HIP(hipMemcpyAsync(r.f_pinned, cnt, FILL_LOG * sizeof(int), hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
  const int used = r.f_pinned[CNT_SLOTS];
  const int overflow = r.f_pinned[CNT_OVERFLOW];
  // sub-iterations with samples (the reference's loop ends at the first
  // round without alive tiles or samples, cache_renderer.cu:1017-1018)
  int subiter = 0;
  int64_t n_eval = 0;
  while (subiter < it && r.f_pinned[FILL_LOG + 2 * subiter] > 0 && r.f_pinned[FILL_LOG + 2 * subiter + 1] > 0) {
    n_eval += r.f_pinned[FILL_LOG + 2 * subiter + 1];
    subiter++;
  }
  for (size_t i = 0; i < hs.size(); i++) {
    const double nt = i < (size_t)subiter ? r.f_pinned[FILL_LOG + 2 * i] : 0.0;
    const double ns = i < (size_t)subiter ? r.f_pinned[FILL_LOG + 2 * i + 1] : 0.0;
    P.work(hs[i].list, n_tiles, n_tiles * 36.0, 0);
    P.work(hs[i].sample, ns, nt * 64.0 * (2 * sizeof(FillRay) + 8) + 20.0 * ns, 0);
    P.work(hs[i].net, ns, ns * (ALG_GATHER_B + (split ? 24.0 : 12.0) + 32.0),
           ns * (ALG_DENSITY_FLOP + (split ? ALG_FIRST_HEAD_FLOP : 0.0)));
    P.work(hs[i].accum, ns, ns * 12.0 + nt * 64.0 * 56, 0);
    P.work(hs[i].alloc, nt, nt * 40.0, 0);
    P.work(hs[i].write, ns, ns * 36.0 + nt * 64.0 * 48, 0);
  }
  r.last_fill_iters = subiter;
  C.n_used = std::min(used, C.capacity);
  C.initialized = true;
  P.collect();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->n_bricks_set = C.n_used;
    stats->n_bricks_reserved = C.capacity;
    stats->overflow = overflow;
    stats->n_subiters = subiter;
    stats->n_samples_evaluated = n_eval;
    stats->samples_ppx = (float)n_eval / (float)((int64_t)W * H);
  }
}

// ------------------------------------------------------------------ render
// sample_cache_performance_kernel (cache_performance_renderer.cu:22-177):
// march, reproject every occupied sample into the cache frustum and
// classify it as resample (outside / beyond the known range), cache hit
// (set brick, nearest-voxel density > 0) or skip.
// The march collects up to n_steps occupied samples per ray (their t0 in
// t_scr[k * n_alive + i]); a second pass classifies them, so the reprojection
// and its three dependent loads stay out of the divergent march loop.  The
// reference stops a launch when either stream reaches n_steps samples; here
// the launch holds n_steps samples in total.  That only moves sub-iteration
// boundaries: every ray still composites the same samples in the same order.
// render_mode ONLY_CACHE / ONLY_RESAMPLED drops the samples of the other
// class (cache_performance_renderer.cu:117-118, 125-126).
// dbg_codes (optional, inspection only): the class of every occupied sample
// of the ray in march order (0 skip, 1 resample, 2 cache), the first dbg_k
// at dbg_codes[pixel * dbg_k ..], dbg_n[pixel] counting them across launches.
static constexpr int CLS_EMIT = 16;  // samples whose t0 the emission loads together
__global__ __launch_bounds__(256) void k_cache_march(int* __restrict__ ctl, int par, int trip_cutoff, MarchGrid mg,
                                                     ContractParams con, const uint32_t* __restrict__ coarse_g,
                                                     CacheView cv, RayDyn* __restrict__ rays, int4* __restrict__ entry,
                                                     int* __restrict__ alive_flag, float* __restrict__ t_scr,
                                                     float4* __restrict__ s_pos, float* __restrict__ s_t0,
                                                     float4* __restrict__ c_rec,
                                                     float* __restrict__ c_t0, const int* __restrict__ pix_in,
                                                     uint4* __restrict__ ray_sh, float4* __restrict__ ray_od,
                                                     uint64_t* __restrict__ order_out,
                                                     const RayState* __restrict__ rconst, int render_mode,
                                                     int8_t* __restrict__ dbg_codes, int dbg_k,
                                                     int* __restrict__ dbg_n) {
  const int n_alive = ctl[CTL_ALIVE + par];
  if ((int)(blockIdx.x * blockDim.x) >= n_alive) return;  // whole block past the alive rays
  const int n_steps = ctl[CTL_NSTEPS];
  int* cnt_r = ctl + CTL_R;
  int* cnt_c = ctl + CTL_C;
  int* step_counter = ctl + CTL_BINS;
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n_alive;
  int cnt = 0, bins = 0;
  uint64_t res_mask = 0, cache_mask = 0;  // class of march sample k
  bool alive = false;
  RayState r{};
  V3 oc = {0, 0, 0}, dc = {0, 0, 0};
  int px = 0;
  if (valid) {
    px = pix_in[i];
    const RayDyn dy = rays[i];
    r = rconst[px];
    r.t = dy.t;
    r.j = dy.j;
    if (r.j == 0) {
      // first launch of the ray: its per-pixel tables (the colour MLP's SH
      // input depends on the direction only; split models also read the ray)
      uint4 lo, hi;
      const V3 d01 = unit_to_01(V3{r.dx, r.dy, r.dz});
      sh4_all(d01.x, d01.y, d01.z, lo, hi);
      ray_sh[2 * (int64_t)px] = lo;
      ray_sh[2 * (int64_t)px + 1] = hi;
      if (ray_od) {
        ray_od[2 * (int64_t)px] = make_float4(r.ox, r.oy, r.oz, 0.0f);
        ray_od[2 * (int64_t)px + 1] = make_float4(r.dx, r.dy, r.dz, 0.0f);
      }
    }
    alive = true;
    MarchRay mr = march_ray(r, mg.tab);
    int trips = 0;
    while (cnt < n_steps && trips < trip_cutoff) {
      trips++;
      float t0;
      const int res = march_trip(mg, coarse, mr, t0, bins);
      if (res == TRIP_END) { alive = false; break; }
      if (res == TRIP_SAMPLE) {
        t_scr[(int64_t)cnt * n_alive + i] = t0;
        cnt++;
      }
    }
    r.t = mr.t1;
    r.j = mr.j;
    // classification (sample_cache_performance_kernel, cache_performance_renderer.cu:96-160)
    oc = transform(cv.cam.world2cam, mr.o.x, mr.o.y, mr.o.z, 1.0f);
    dc = transform(cv.cam.world2cam, mr.d.x, mr.d.y, mr.d.z, 0.0f);
    // CLS samples per step; each step issues all its loads (t0, known range,
    // brick slot, then voxel density) before using them, from clamped
    // addresses whose results are selected away where they do not apply
    constexpr int CLS = 4;
    const float* dens = cv.lat8 ? &cv.pool8[0].density : &cv.pool[0].density;
    const int vstride = cv.lat8 ? (int)(sizeof(Voxel8) / 4) : (int)(sizeof(Voxel) / 4);
    for (int k0 = 0; k0 < cnt; k0 += CLS) {
      float t0v[CLS], tf[CLS], known[CLS], dv[CLS];
      V3 f[CLS];
      bool in_fov[CLS], inside[CLS];
      int slot[CLS], loc[CLS];
#pragma unroll
      for (int u = 0; u < CLS; u++) t0v[u] = t_scr[(int64_t)min(k0 + u, cnt - 1) * n_alive + i];
#pragma unroll
      for (int u = 0; u < CLS; u++) {
        const float t0 = t0v[u];
        const float t1 = t0 + step_dt(mg.st, t0);
        const float tm = (t0 + t1) * 0.5f;
        f[u] = cam2froxel(oc + tm * dc, cv.cam, cv.st, tf[u]);
        in_fov[u] = f[u].x >= 0.0f && f[u].x < (float)cv.fx && f[u].y >= 0.0f && f[u].y < (float)cv.fy;
        inside[u] = in_fov[u] && f[u].z >= 0.0f && f[u].z < (float)cv.fz;
        const int vx = inside[u] ? (int)floorf(f[u].x) : 0, vy = inside[u] ? (int)floorf(f[u].y) : 0;
        const int vz = inside[u] ? (int)floorf(f[u].z) : 0;
        const int64_t kidx = in_fov[u] ? (int64_t)(int)f[u].y * cv.fx + (int)f[u].x : 0;
        known[u] = cv.known[kidx];
        slot[u] = cv.table[((int64_t)(vz >> 3) * cv.by + (vy >> 3)) * cv.bx + (vx >> 3)];
        loc[u] = brick_voxel(vx, vy, vz);
      }
#pragma unroll
      for (int u = 0; u < CLS; u++) {
        const int64_t v = slot[u] >= 0 ? (int64_t)slot[u] * BRICK_VOXELS + loc[u] : 0;
        dv[u] = dens[v * vstride];
      }
#pragma unroll
      for (int u = 0; u < CLS; u++) {
        if (k0 + u >= cnt) break;
        const bool unsampled = !in_fov[u] || tf[u] < cv.st.near_ || tf[u] > known[u];
        if (!inside[u] || unsampled) {
          if (render_mode != FVC_RENDER_ONLY_CACHE) res_mask |= 1ull << (k0 + u);
        } else if (render_mode != FVC_RENDER_ONLY_RESAMPLED && slot[u] >= 0 && dv[u] > 0.0f) {
          cache_mask |= 1ull << (k0 + u);
        }
      }
    }
  }
  const int nr = __popcll(res_mask), nc = __popcll(cache_mask);
  int slot[3];
  block_reserve<3>({nr, nc, bins}, {cnt_r, cnt_c, step_counter}, slot);
  const int off_r = slot[0], off_c = slot[1];
  if (valid) {
    const V3 o = {r.ox, r.oy, r.oz}, d = {r.dx, r.dy, r.dz};
    const float pxf = __builtin_bit_cast(float, px);
    uint64_t order = 0;  // bit q set: the q-th kept sample is a cache sample
    int kr = 0, kc = 0;
    const uint64_t kept = res_mask | cache_mask;
    for (int k0 = 0; k0 < cnt; k0 += CLS_EMIT) {
      if (((kept >> k0) & ((1ull << CLS_EMIT) - 1ull)) == 0ull) continue;
      float t0v[CLS_EMIT];
#pragma unroll
      for (int u = 0; u < CLS_EMIT; u++) t0v[u] = t_scr[(int64_t)min(k0 + u, cnt - 1) * n_alive + i];
#pragma unroll
      for (int u = 0; u < CLS_EMIT; u++) {
        const int k = k0 + u;
        if (k >= cnt) break;
        const bool is_r = (res_mask >> k) & 1ull, is_c = (cache_mask >> k) & 1ull;
        if (!is_r && !is_c) continue;
        const float t0 = t0v[u];
        const float tm2 = t0 + step_dt(mg.st, t0) * 0.5f;
        if (is_r) {
          // generateNetworkInput_resample_perf_kernel (198-240): t_mid = t0 + dt/2
          const V3 p = contract(o + tm2 * d, con);
          const int64_t q = (int64_t)off_r + kr;
          s_pos[q] = make_float4(p.x, p.y, p.z, pxf);
          s_t0[q] = t0;
        } else {
          // generateNetworkInput_preparation_perf_kernel (242-313)
          float tf;
          const V3 f = cam2froxel(oc + tm2 * dc, cv.cam, cv.st, tf);
          const int64_t q = (int64_t)off_c + kc;
          c_rec[q] = make_float4(f.x, f.y, f.z, __builtin_bit_cast(float, px));
          c_t0[q] = t0;
          order |= 1ull << (kr + kc);
        }
        kr += is_r ? 1 : 0;
        kc += is_c ? 1 : 0;
      }
    }
    rays[i] = RayDyn{r.t, r.j};
    entry[i] = make_int4(off_r, nr, off_c, nc);
    order_out[i] = order;
    alive_flag[i] = alive;
    if (dbg_codes) {
      const int base = dbg_n[px];
      for (int k = 0; k < cnt; k++)
        if (base + k < dbg_k)
          dbg_codes[(int64_t)px * dbg_k + base + k] =
              (int8_t)(((res_mask >> k) & 1ull) ? 1 : (((cache_mask >> k) & 1ull) ? 2 : 0));
      dbg_n[px] = base + cnt;
    }
  }
}

// Corners outside the frustum or in unset bricks read this zero voxel.
__device__ Voxel k_zero_voxel;  // zero-initialised, never written

// 8 E4M3 codes (latent i in byte i % 4 of word i / 4) times the brick scale
template <int B> __device__ __forceinline__ float fp8_byte(uint32_t w) { return __builtin_amdgcn_cvt_f32_fp8((int)w, B); }
__device__ __forceinline__ void fp8x8_decode(uint32_t w0, uint32_t w1, float s, float* out) {
  out[0] = fp8_byte<0>(w0) * s; out[1] = fp8_byte<1>(w0) * s; out[2] = fp8_byte<2>(w0) * s; out[3] = fp8_byte<3>(w0) * s;
  out[4] = fp8_byte<0>(w1) * s; out[5] = fp8_byte<1>(w1) * s; out[6] = fp8_byte<2>(w1) * s; out[7] = fp8_byte<3>(w1) * s;
}

// The trilinear lookup of one cache sample by one lane: its 8 corners
// (corner c = 4x + 2y + z; outside the frustum or in unset bricks the zero
// voxel) summed in the reference's order ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7))
// (cache_performance_renderer.cu:736-787): wd = sum w*isset, dd = sum
// w*sigma, sums[k] = sum w*l_k.  rec = (froxel x, y, z, -).
// applyInterpolationFunction (raymarch_common.h:393-402) of one fraction:
// smoothstep x*x*(3 - 2x), smootherstep x*x*x*(x*(6x - 15) + 10)
// (helper_math_extension.h:146-170, the same evaluation order), nearest =
// round half to even.
__device__ __forceinline__ float interp_weight(float t, int fn) {
  if (fn == FVC_FN_SMOOTHSTEP) return (t * t) * (3.0f - 2.0f * t);
  if (fn == FVC_FN_SMOOTHERSTEP) return ((t * t) * t) * (t * (6.0f * t - 15.0f) + 10.0f);
  if (fn == FVC_FN_NEAREST) return __builtin_rintf(t);
  return t;
}
// use_inter_brick_interpolation = false: the coordinate clamped into
// [8b + 1/2, 8b + 7.5] of its brick b (froxel2data, raymarch_common.h:341-357)
__device__ __forceinline__ float clamp_to_brick(float x) {
  const float b = floorf(x * 0.125f) * 8.0f;
  return b + fminf_(fmaxf_(x - b, 0.5f), 7.5f);
}

template <bool LAT8>
__device__ __forceinline__ void lookup_sample(const CacheView& cv, float4 rec, float& wd, float& dd,
                                              float* sums) {
  if (!cv.inter_brick) {
    rec.x = clamp_to_brick(rec.x);
    rec.y = clamp_to_brick(rec.y);
    rec.z = clamp_to_brick(rec.z);
  }
  const float fcx = rec.x - 0.5f, fcy = rec.y - 0.5f, fcz = rec.z - 0.5f;
  const float ix = floorf(fcx), iy = floorf(fcy), iz = floorf(fcz);
  float lx = fcx - ix, ly = fcy - iy, lz = fcz - iz;
  if (cv.interp_fn != FVC_FN_LINEAR) {
    lx = interp_weight(lx, cv.interp_fn);
    ly = interp_weight(ly, cv.interp_fn);
    lz = interp_weight(lz, cv.interp_fn);
  }
  // weight factors (1 - off)(1 - 2l) + l of offsets 0 and 1
  const float wxs[2] = {(1.0f - 2.0f * lx) + lx, lx}, wys[2] = {(1.0f - 2.0f * ly) + ly, ly};
  const float wzs[2] = {(1.0f - 2.0f * lz) + lz, lz};
  uint4 lat[8];
  float2 dis[8];
  float scl[8];
  const Voxel* vp[8];  // FP16 pool: corner voxels (the zero voxel outside)
  bool vok[8];
  // integer cell of the corner (0, 0, 0); corner coordinates are exact
  // small integers, so the frustum test is an unsigned compare
  const int ixi = (int)ix, iyi = (int)iy, izi = (int)iz;
  // brick slots: the corners share the brick of corner (0, 0, 0) unless an
  // offset crosses a brick face, so only lanes whose corner lies in another
  // brick load its slot (masked lanes cost no cache access)
  const uint32_t bx0 = (uint32_t)ixi >> 3, by0 = (uint32_t)iyi >> 3, bz0 = (uint32_t)izi >> 3;
  const bool crx = (ixi & 7) == 7, cry = (iyi & 7) == 7, crz = (izi & 7) == 7;
  const uint32_t b0 = (bz0 * (uint32_t)cv.by + by0) * (uint32_t)cv.bx + bx0;
  // corner (0, 0, 0) lies in the frustum whenever any corner does; its
  // brick index is only read then (cell -1 gives a huge unsigned index)
  const bool in0 = (uint32_t)ixi < (uint32_t)cv.fx && (uint32_t)iyi < (uint32_t)cv.fy &&
                   (uint32_t)izi < (uint32_t)cv.fz;
  const int slot0 = cv.table[in0 ? b0 : 0u];
  // every slot load is issued before any voxel load waits for one
  int slots[8];
  bool ins[8], own[8];  // own: the corner's brick is not corner 0's
  uint32_t locs[8];
  // FP8 latents: the brick scale of corner 0, shared like its slot
  const float scale0 = LAT8 ? cv.scales[in0 && slot0 >= 0 ? (uint32_t)slot0 : 0u] : 0.0f;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int ox = (c >> 2) & 1, oy = (c >> 1) & 1, oz = c & 1;
    const uint32_t cx = (uint32_t)(ixi + ox), cy = (uint32_t)(iyi + oy), cz = (uint32_t)(izi + oz);
    const bool inside = cx < (uint32_t)cv.fx && cy < (uint32_t)cv.fy && cz < (uint32_t)cv.fz;
    const uint32_t vx = inside ? cx : 0u, vy = inside ? cy : 0u, vz = inside ? cz : 0u;
    const bool other = (ox && crx) || (oy && cry) || (oz && crz) || !in0;
    slots[c] = slot0;
    own[c] = c != 0 && other && inside;
    if (own[c]) slots[c] = cv.table[((vz >> 3) * (uint32_t)cv.by + (vy >> 3)) * (uint32_t)cv.bx + (vx >> 3)];
    ins[c] = inside;
    locs[c] = (uint32_t)brick_voxel((int)vx, (int)vy, (int)vz);
  }
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int slot = slots[c];
    const bool ok = ins[c] && slot >= 0;
    const size_t vi = (size_t)(uint32_t)slot * BRICK_VOXELS + locs[c];
    if (LAT8) {
      const uint4 q = *reinterpret_cast<const uint4*>(ok ? (const void*)(cv.pool8 + vi) : (const void*)&k_zero_voxel);
      lat[c] = make_uint4(q.x, q.y, 0u, 0u);
      dis[c] = make_float2(__builtin_bit_cast(float, q.z), __builtin_bit_cast(float, q.w));
      float sc = scale0;
      if (own[c] && ok) sc = cv.scales[(uint32_t)slot];
      scl[c] = ok ? sc : 0.0f;
    } else {
      vok[c] = ok;
      vp[c] = ok ? cv.pool + vi : &k_zero_voxel;
      lat[c] = *reinterpret_cast<const uint4*>(vp[c]->lat);
      scl[c] = 0.0f;
    }
  }
  if (!LAT8) {
    // density and isset: corner c (z offset 0) with its znext, which is
    // corner c + 1 when both are voxels of one brick column; otherwise
    // corner c + 1 loads its own (only those lanes access memory)
    float4 dh[4];
    float2 d1[4];
    bool shared[4];
#pragma unroll
    for (int p = 0; p < 4; p++) {
      const int c = 2 * p;
      dh[p] = *reinterpret_cast<const float4*>(&vp[c]->density);
      shared[p] = vok[c] && vok[c + 1] && !crz;
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
      d1[p] = make_float2(0.0f, 0.0f);
      if (!shared[p]) d1[p] = *reinterpret_cast<const float2*>(&vp[2 * p + 1]->density);
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
      dis[2 * p] = make_float2(dh[p].x, dh[p].y);
      dis[2 * p + 1] = shared[p] ? make_float2(dh[p].z, dh[p].w) : d1[p];
    }
  }
  // per-corner products, then the reference's summation tree.  FP16
  // latents: w * l as fma(l, w, -0), the exact product (signed zeros
  // included), which takes the half operand directly (v_fma_mix_f32); the
  // -0 is opaque so that the fma is not folded back into a multiply.
  const float nz = opaque_f(-0.0f);
  float cw[8], cd[8], cl[8][8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const float w = (wxs[(c >> 2) & 1] * wys[(c >> 1) & 1]) * wzs[c & 1];
    cw[c] = w * dis[c].y;
    cd[c] = w * dis[c].x;
    if (LAT8) {
      float lv[8];
      fp8x8_decode(lat[c].x, lat[c].y, scl[c], lv);
#pragma unroll
      for (int k = 0; k < 8; k++) cl[k][c] = w * lv[k];
    } else {
      const uint32_t lw[4] = {lat[c].x, lat[c].y, lat[c].z, lat[c].w};
#pragma unroll
      for (int k = 0; k < 8; k++) cl[k][c] = __builtin_fmaf(h2f((uint16_t)(lw[k >> 1] >> (16 * (k & 1)))), w, nz);
    }
  }
  auto tree = [](const float* v) { return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7])); };
  wd = tree(cw);
  dd = tree(cd);
#pragma unroll
  for (int k = 0; k < 8; k++) sums[k] = tree(cl[k]);
}

// Head inputs of one cache sample from its lookup: the 8 latents
// Sum w*l / Sum w*isset and the density row log(max(Sum w*sigma, 1e-6)),
// rounded to half (generateNetworkInput_cache_shfl_perf_kernel_cg,
// cache_performance_renderer.cu:736-787).  di (DensityIntermediates): latent
// 0 holds the density, latent 0 = log(Sum w*l0) * W / W with W = Sum w*isset
// (770-779, no clamp), and it is also the density row.  Sum w*isset >= 1/8
// for every interpolation function but nearest (the voxel holding the sample
// is set and weighs at least 1/2 per axis), so one correctly rounded
// reciprocal serves the 8 divisions; nearest weights are 0 or 1 and divide
// as IEEE division does (0/0 included).
template <bool LAT8>
__device__ __forceinline__ void cache_head_inputs(const CacheView& cv, float4 rec, int di, _Float16* L,
                                                  _Float16& logsig) {
  float wd, dd, sums[8];
  lookup_sample<LAT8>(cv, rec, wd, dd, sums);
  if (cv.interp_fn == FVC_FN_NEAREST) {
#pragma unroll
    for (int k = 0; k < 8; k++) L[k] = (_Float16)(sums[k] / wd);
    if (di) L[0] = (_Float16)((logf(sums[0]) * wd) / wd);
  } else {
    // the quotient is rounded to float, then to half, as the reference's
    // __float2half(a / b) does (f16_of: not fused with div_recip's last fma)
    const float rw = 1.0f / wd;
#pragma unroll
    for (int k = 0; k < 8; k++) L[k] = f16_of(div_recip(sums[k], wd, rw));
    if (di) L[0] = f16_of(div_recip(logf(sums[0]) * wd, wd, rw));
  }
  logsig = di ? L[0] : (_Float16)logf(fmaxf_(dd, 1e-6f));
}

// One lane per sample: a wave takes 64 cache samples.  Each lane fetches
// its sample's 8 corners (corner c = 4x + 2y + z, outside the frustum or in
// unset bricks the zero voxel), sums them in the reference's order
// ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)) (cache_performance_renderer.cu:736-
// 787), forms the colour-MLP input row -- latents sum w*l / sum w*isset, SH of
// its pixel's ray, the constant 1.0 padding (nerf_network.h:248-253) -- and
// writes it to LDS in the MFMA operand order of FRAG_H1 (lane group g of a
// tile reads the 8 halves of its K slots); the wave then evaluates its four
// 16-sample tiles.  The corner arithmetic is done once per sample instead of
// once per lane of a 4-lane group.  LAT8: the pool holds FP8 latents (Voxel8)
// decoded as code * brick scale.
static const int HEAD_REC = 5;  // uint4 per sample record in LDS: 4 operand groups + log sigma (80 bytes)

// The colour head's first 16 inputs are the density-MLP output rows in the
// layout of the model (sc.lat_lo, sc.lat_hi): latents in rows [lo, lo + n),
// the constant 1.0 elsewhere; di: DensityIntermediates (the reference's plain
// models, lo = 0).
template <int P, bool LAT8>
__global__ __launch_bounds__(256) void k_cache_head(const float4* __restrict__ c_rec, const uint4* __restrict__ ray_sh,
                                                    int64_t n, CacheView cv,
                                                    const typename Mx<P>::T* __restrict__ frags, MlpScales sc,
                                                    int di, uint2* __restrict__ c_out, const int* __restrict__ n_dev) {
  // n_dev: sample count written by the marcher (n is then its bound)
  if (n_dev) n = *n_dev;
  if ((int64_t)blockIdx.x * 256 >= n) return;
  typedef typename Mx<P>::T T;
  __shared__ T wl[6 * 64];  // FRAG_H1 (4), FRAG_H2 (2)
  __shared__ uint4 rec_lds[4][64 * HEAD_REC];
  stage_frags(frags + FRAG_H1 * 64, wl, 6);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int n32 = (int)n;  // the stream holds fewer than 2^31 / 4 samples
  const int n_groups = (n32 + 63) / 64;
  uint4* recw = rec_lds[wave];
  const uint32_t one2 = 0x3c003c00u;  // two fp16 1.0
  const int gstride = gridDim.x * 4;
  // the record of the wave's next group is loaded one iteration ahead, so
  // the corner loads wait for one load less
  int grp = blockIdx.x * 4 + wave;
  float4 rec_next = c_rec[min(grp * 64 + lane, n32 - 1)];
  for (; grp < n_groups; grp += gstride) {
    const int li = opaque(lane);
    const float4 rec = rec_next;
    rec_next = c_rec[min((grp + gstride) * 64 + lane, n32 - 1)];
    const int px = __builtin_bit_cast(int, rec.w);
    const uint4 sh_lo = ray_sh[2 * (int64_t)px], sh_hi = ray_sh[2 * (int64_t)px + 1];
    _Float16 L[8], logsig;
    cache_head_inputs<LAT8>(cv, rec, di, L, logsig);
    auto pk = [](_Float16 a, _Float16 b) {
      return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
    };
    const _Float16 h1 = (_Float16)1.0f;
    // operand groups: base-output rows 4g..4g+3 (latents in rows [lo, lo + n),
    // the rest the padding 1.0; lo = 1 with n = 8, or lo = 0 with n <= 8) and
    // SH coefficients 4g..4g+3
    const int lo = sc.lat_lo, hi = sc.lat_hi;
    _Float16 R[9];
    R[0] = lo ? h1 : L[0];
#pragma unroll
    for (int r = 1; r < 8; r++) R[r] = r < hi ? (lo ? L[r - 1] : L[r]) : h1;
    R[8] = hi > 8 ? L[7] : h1;
    uint4* r5 = recw + lane * HEAD_REC;
    r5[0] = make_uint4(pk(R[0], R[1]), pk(R[2], R[3]), sh_lo.x, sh_lo.y);
    r5[1] = make_uint4(pk(R[4], R[5]), pk(R[6], R[7]), sh_lo.z, sh_lo.w);
    r5[2] = make_uint4(pk(R[8], h1), one2, sh_hi.x, sh_hi.y);
    r5[3] = make_uint4(one2, one2, sh_hi.z, sh_hi.w);
    r5[4] = make_uint4((uint32_t)__builtin_bit_cast(uint16_t, logsig), 0u, 0u, 0u);
    wave_sync();
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const uint4* rt = recw + (16 * t + col) * HEAD_REC;
      const uint4 b = rt[g];
      T x;
      if constexpr (P == FVC_PREC_FP16) {
        x = __builtin_bit_cast(T, b);
      } else {
        const uint32_t w4[4] = {b.x, b.y, b.z, b.w};
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = h2f((uint16_t)(w4[k >> 1] >> (16 * (k & 1))));
        x = Mx<P>::pack(v, sc.inv_act);
      }
      const f4 o = head_mlp<P>(wl, wl + 4 * 64, li, x, sc);
      const int st = grp * 64 + 16 * t + col;
      if (g == 0 && st < n32) {
        uint2 v;
        v.x = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)o[0]) |
              ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)o[1]) << 16);
        v.y = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)o[2]) | (rt[4].x << 16);
        c_out[st] = v;
      }
    }
    wave_sync();
  }
}

// Split models: the lookup of k_cache_head (FP16 latents, one lane per
// sample) with the custom colour head (generateNetworkInput_cache_padded_kernel
// / _shfl_perf_kernel_cg + inferenceHead, cache_performance_renderer.cu:
// 323-410, 678-787).  Per sample the midpoint t0 + dt/2 on its pixel's ray
// gives the position of the frequency encoding and the initial view
// direction (normalize(point - cache camera position), 292-312).  di
// (DensityIntermediates): latent 0 holds the density, so latent 0 =
// log(sum w l_0) * W / W with W = sum w isset (770-779, no clamp) and it is
// also the sample's log density.  The 64 head inputs go to LDS as the two
// operand blocks (VD_HA, VD_HB order) of each lane group.
static const int HEAD_REC_VD = 9;  // uint4 per sample record: block A (4 groups), block B (4 groups), log sigma

__global__ __launch_bounds__(256) void k_cache_head_vd(const float4* __restrict__ c_rec, const float* __restrict__ c_t0,
                                                       const uint4* __restrict__ ray_sh,
                                                       const float4* __restrict__ ray_od, int64_t n, CacheView cv,
                                                       fvc_stepsize_info st, ContractParams con,
                                                       const h8* __restrict__ frags, int di, int freq_on,
                                                       uint2* __restrict__ c_out, const int* __restrict__ n_dev) {
  if (n_dev) n = *n_dev;
  if ((int64_t)blockIdx.x * 256 >= n) return;
  __shared__ h8 wl[10 * 64];  // VD_HA (4), VD_HB (4), VD_HO (2)
  __shared__ uint4 rec_lds[4][64 * HEAD_REC_VD];
  stage_frags(frags + VD_HA * 64, wl, 10);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int n32 = (int)n;
  const int n_groups = (n32 + 63) / 64;
  const V3 cam_pos = {cv.cam.cam2world.m0.w, cv.cam.cam2world.m1.w, cv.cam.cam2world.m2.w};
  uint4* recw = rec_lds[wave];
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  auto pk = [](_Float16 a, _Float16 b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
  };
  const int gstride = gridDim.x * 4;
  // record and t0 of the wave's next group loaded one iteration ahead
  int grp = blockIdx.x * 4 + wave;
  float4 rec_next = c_rec[min(grp * 64 + lane, n32 - 1)];
  float t0_next = c_t0[min(grp * 64 + lane, n32 - 1)];
  for (; grp < n_groups; grp += gstride) {
    const int li = opaque(lane);
    const float4 rec = rec_next;
    const float t0 = t0_next;
    const int sn = min((grp + gstride) * 64 + lane, n32 - 1);
    rec_next = c_rec[sn];
    t0_next = c_t0[sn];
    const int px = __builtin_bit_cast(int, rec.w);
    const uint4 sh_lo = ray_sh[2 * (int64_t)px], sh_hi = ray_sh[2 * (int64_t)px + 1];
    const float4 ro = ray_od[2 * (int64_t)px], rd = ray_od[2 * (int64_t)px + 1];
    _Float16 L[8], logsig;
    cache_head_inputs<false>(cv, rec, di, L, logsig);
    // the sample's midpoint (generateNetworkInput_preparation_perf_kernel, 285-291)
    const V3 o = {ro.x, ro.y, ro.z}, d = {rd.x, rd.y, rd.z};
    const float tm = t0 + step_dt(st, t0) * 0.5f;
    const V3 wp = o + tm * d;
    const V3 p = contract(wp, con);
    const V3 i01 = unit_to_01(normalize(wp - cam_pos));
    uint4 ilo, ihi;
    sh4_all(i01.x, i01.y, i01.z, ilo, ihi);
    // SH differences: the half coefficients subtracted in half (shCoeffDiffChunk4)
    const uint32_t dw[8] = {sh_lo.x, sh_lo.y, sh_lo.z, sh_lo.w, sh_hi.x, sh_hi.y, sh_hi.z, sh_hi.w};
    const uint32_t iw[8] = {ilo.x, ilo.y, ilo.z, ilo.w, ihi.x, ihi.y, ihi.z, ihi.w};
    uint32_t df[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
      df[k] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(h2, dw[k]) - __builtin_bit_cast(h2, iw[k]));
    _Float16 fq[6][4];
#pragma unroll
    for (int q = 0; q < 6; q++) freq_quad(p, q, freq_on != 0, fq[q]);
    uint4* r9 = recw + lane * HEAD_REC_VD;
    // block A: latents 4g..4g+3 (g < 2) or frequency quad g + 2 (g >= 2); SH 4g..4g+3
    r9[0] = make_uint4(pk(L[0], L[1]), pk(L[2], L[3]), dw[0], dw[1]);
    r9[1] = make_uint4(pk(L[4], L[5]), pk(L[6], L[7]), dw[2], dw[3]);
    r9[2] = make_uint4(pk(fq[4][0], fq[4][1]), pk(fq[4][2], fq[4][3]), dw[4], dw[5]);
    r9[3] = make_uint4(pk(fq[5][0], fq[5][1]), pk(fq[5][2], fq[5][3]), dw[6], dw[7]);
    // block B: SH differences 4g..4g+3, frequency quad g
#pragma unroll
    for (int q = 0; q < 4; q++)
      r9[4 + q] = make_uint4(df[2 * q], df[2 * q + 1], pk(fq[q][0], fq[q][1]), pk(fq[q][2], fq[q][3]));
    r9[8] = make_uint4((uint32_t)__builtin_bit_cast(uint16_t, logsig), 0u, 0u, 0u);
    wave_sync();
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const uint4* rt = recw + (16 * t + col) * HEAD_REC_VD;
      const h8 xa = __builtin_bit_cast(h8, rt[g]), xb = __builtin_bit_cast(h8, rt[4 + g]);
      const f4 o4 = head_vd_mlp(wl, wl + 4 * 64, wl + 8 * 64, li, xa, xb);
      const int st2 = grp * 64 + 16 * t + col;
      if (g == 0 && st2 < n32) {
        uint2 v;
        v.x = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)o4[0]) |
              ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)o4[1]) << 16);
        v.y = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)o4[2]) | (rt[8].x << 16);
        c_out[st2] = v;
      }
    }
    wave_sync();
  }
}

// accumulate_cache_perf_kernel (cache_performance_renderer.cu:789-908): merge
// the resample and cache streams of a ray by t0 (resample first iff t0_r <
// t0_c) and composite; then compact the alive rays.
__global__ __launch_bounds__(256) void k_cache_composite(
    int* __restrict__ ctl, int par, fvc_stepsize_info st, float alpha_thre, const RayDyn* __restrict__ rays_in,
    const int* __restrict__ pix_in, const float4* __restrict__ res_in, const int4* __restrict__ entry,
    const uint64_t* __restrict__ order, const int* __restrict__ alive_flag, const uint2* __restrict__ s_out,
    const float* __restrict__ s_t0, const uint2* __restrict__ c_out, const float* __restrict__ c_t0,
    RayDyn* __restrict__ rays_out, int* __restrict__ pix_out, float4* __restrict__ res_out,
    float4* __restrict__ results, int2* __restrict__ dbg_comp) {
  // dbg_comp (optional, inspection only): per pixel the resampled and cached
  // samples composited, up to and including the one that ends the ray
  const int n_alive = ctl[CTL_ALIVE + par];
  if ((int)(blockIdx.x * blockDim.x) >= n_alive) return;
  int* counter = ctl + CTL_ALIVE + (par ^ 1);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n_alive;
  bool keep = false;
  int px = 0;
  float4 res = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (valid) {
    res = res_in[i];
    const int4 e = entry[i];
    // loaded with the entry: they are only needed after the samples
    const int ray_alive = alive_flag[i];
    px = pix_in[i];
    const uint64_t ord = order[i];
    const int n = e.y + e.w;
    bool visible = true;  // transmittance still above the early-stop threshold
    int kr = 0, kc = 0;
    int n_done = 0;       // samples composited
    // the two streams interleaved in march (= t0) order, batch by batch
    for (int k0 = 0; k0 < n && visible; k0 += COMPOSITE_BATCH) {
      uint2 v[COMPOSITE_BATCH];
      float t[COMPOSITE_BATCH];
#pragma unroll
      for (int u = 0; u < COMPOSITE_BATCH; u++) {
        if (k0 + u < n) {
          const bool from_cache = (ord >> (k0 + u)) & 1u;
          const int q = from_cache ? e.z + kc : e.x + kr;
          kc += from_cache ? 1 : 0;
          kr += from_cache ? 0 : 1;
          v[u] = (from_cache ? c_out : s_out)[q];
          t[u] = (from_cache ? c_t0 : s_t0)[q];
        }
      }
#pragma unroll
      for (int u = 0; u < COMPOSITE_BATCH; u++)
        if (visible && k0 + u < n) {
          visible = composite_sample(res, v[u], t[u], st, alpha_thre);
          n_done++;
        }
    }
    keep = visible && ray_alive != 0;
    if (!keep) results[px] = res;
    if (dbg_comp) {
      int cc = 0;
      for (int k = 0; k < n_done; k++) cc += (int)((ord >> k) & 1u);
      const int2 d = dbg_comp[px];
      dbg_comp[px] = make_int2(d.x + n_done - cc, d.y + cc);
    }
  }
  int slot[1];
  block_reserve<1>({keep ? 1 : 0}, {counter}, slot);
  if (keep) {
    rays_out[slot[0]] = rays_in[i];
    pix_out[slot[0]] = px;
    res_out[slot[0]] = res;
  }
}

void render_cached(Renderer& r, const Views& vs, fvc_render_stats* stats) {
  const Model& m = *r.model;
  check_cache_latents(m);
  const CacheStore& C = r.caches[r.render_idx];
  if (!C.initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
  if (vs.scenes.size() > 1) throw Error(FVC_E_UNSUPPORTED, "a cached batch renders the scene of the renderer's cache");
  prepare_views(r, vs);
  const fvc_raymarch_info& rm = vs.rms[0];
  const double n_px = (double)vs.n_px;
  const int64_t cap = r.sample_capacity;
  r.c_rec.ensure(cap * sizeof(float4));
  r.ray_sh.ensure(vs.n_px * 32);
  if (m.net.split_head) r.ray_od.ensure(vs.n_px * 32);
  r.c_t0.ensure(cap * sizeof(float));
  r.c_out.ensure(cap * sizeof(uint2));
  hipStream_t s = r.stream();
  CacheView cv = cache_view(C);
  cv.interp_fn = r.settings.interpol_function;
  cv.inter_brick = r.settings.use_inter_brick_interpolation ? 1 : 0;
  // DensityIntermediates for the reference's plain models (latent 0 = density)
  const int di_plain = !m.net.split_head && m.net.first_latent_is_density ? 1 : 0;
  const MarchGrid mg = march_grid(r, vs.rms, vs.n);

  Profiler& P = r.prof;
  // graph mode as in render_full
  const bool graph = r.use_graph && r.graph_iters[1] > 0 && !P.enabled && !r.dbg_codes;
  hipEvent_t ev = P.begin(s);
  const int n0 = graph ? (int)vs.n_px : launch_init_rays(r, vs, s);
  P.work(P.end(ev, s, "init_rays"), n_px, n_px * 16 + n0 * 60.0, 0);
  int* ctl = r.counters.as<int>();
  const int64_t target = std::min<int64_t>(vs.target(r.target_batch), cap - 64);
  const int64_t n_bound = std::min<int64_t>(cap - 64, (int64_t)n0 * r.max_steps);  // samples per launch
  const dim3 grid((n0 + 255) / 256);
  const int head_blocks = (int)std::min<int64_t>((n_bound + 255) / 256, 256 * 16);
  // loop trips per ray and launch; a performance knob only (a ray resumes
  // where it stopped), like the reference's step_cutoff
  // (cache_performance_renderer.cu:72)
  const int cutoff = 4096;
  struct Handles { int march, net, head, comp; };
  std::vector<Handles> hs;
  auto iteration = [&](int it) {
      const int cur = it & 1;
      hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, s, ctl, it, target, r.max_steps, cap - 64, 0);
      Handles h;
      ev = P.begin(s);
      hipLaunchKernelGGL(k_cache_march, grid, dim3(256), 0, s, ctl, cur, cutoff, mg, m.con, m.coarse.as<uint32_t>(), cv,
                         r.rays[cur].as<RayDyn>(), r.ray_entry.as<int4>(), r.ray_alive.as<int>(),
                         r.t_scratch.as<float>(), r.s_pos.as<float4>(), r.s_t0.as<float>(), r.c_rec.as<float4>(), r.c_t0.as<float>(), r.ray_pix[cur].as<int>(), r.ray_sh.as<uint4>(),
                         m.net.split_head ? r.ray_od.as<float4>() : nullptr,
                         r.ray_order.as<uint64_t>(), r.ray_const.as<RayState>(), r.render_mode, r.dbg_codes, r.dbg_k,
                         r.dbg_n);
      FVC_HIP(hipGetLastError());
      h.march = P.end(ev, s, "cache_march");
      ev = P.begin(s);
      launch_network_full_pix(m, r.s_pos.as<float4>(), r.ray_sh.as<uint4>(), n_bound, r.s_out.as<uint32_t>(), s,
                              ctl + CTL_R, r.mlp_rows);
      h.net = P.end(ev, s, "network_resample");
      ev = P.begin(s);
#define L2(P, F8)                                                                                                    \
  hipLaunchKernelGGL((k_cache_head<P, F8>), dim3(head_blocks), dim3(256), 0, s, r.c_rec.as<float4>(),                 \
                     r.ray_sh.as<uint4>(), (int64_t)n_bound, cv, m.frags.as<typename Mx<P>::T>(), m.mlp_scales,       \
                     di_plain, r.c_out.as<uint2>(), ctl + CTL_C)
#define L(P)                     \
  if (cv.lat8) { L2(P, true); }  \
  else { L2(P, false); }
      if (m.net.split_head) {
        hipLaunchKernelGGL(k_cache_head_vd, dim3(head_blocks), dim3(256), 0, s, r.c_rec.as<float4>(),
                           r.c_t0.as<float>(), r.ray_sh.as<uint4>(), r.ray_od.as<float4>(), (int64_t)n_bound, cv,
                           rm.stepsize_info, m.con, m.frags.as<h8>(), m.net.first_latent_is_density,
                           m.net.freq_encoding_degree > 0, r.c_out.as<uint2>(), ctl + CTL_C);
      } else {
        switch (m.net.precision) {
          case FVC_PREC_BF16: L(FVC_PREC_BF16); break;
          case FVC_PREC_FP8: L(FVC_PREC_FP8); break;
          case FVC_PREC_BF16X3: L(FVC_PREC_BF16X3); break;
          default: L(FVC_PREC_FP16); break;
        }
      }
#undef L
#undef L2
      FVC_HIP(hipGetLastError());
      h.head = P.end(ev, s, "cache_head");
      ev = P.begin(s);
      hipLaunchKernelGGL(k_cache_composite, grid, dim3(256), 0, s, ctl, cur, rm.stepsize_info, m.scene.alpha_thre,
                         r.rays[cur].as<RayDyn>(), r.ray_pix[cur].as<int>(), r.ray_res[cur].as<float4>(),
                         r.ray_entry.as<int4>(), r.ray_order.as<uint64_t>(), r.ray_alive.as<int>(),
                         r.s_out.as<uint2>(), r.s_t0.as<float>(), r.c_out.as<uint2>(), r.c_t0.as<float>(),
                         r.rays[cur ^ 1].as<RayDyn>(), r.ray_pix[cur ^ 1].as<int>(),
                         r.ray_res[cur ^ 1].as<float4>(), r.results.as<float4>(), r.dbg_comp);
      FVC_HIP(hipGetLastError());
      h.comp = P.end(ev, s, "cache_composite");
      hs.push_back(h);
  };
  int it = 0;
  bool done = n0 == 0;
  if (graph) {
    const int k = std::min(r.graph_iters[1] + 1, MAX_SUBITER);
    graph_run(r, 1, s, [&] {
      launch_init_rays_async(r, vs, s);
      for (int i = 0; i < k; i++) iteration(i);
    });
    it = k;
    done = read_counter(r, CTL_ALIVE + (it & 1), s) == 0;
  }
  while (!done && it < MAX_SUBITER) {
    const int burst = next_burst(r, it);
    for (int b = 0; b < burst && it < MAX_SUBITER; b++, it++) iteration(it);
    done = read_counter(r, CTL_ALIVE + (it & 1), s) == 0;
  }
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, s, ctl, it, target, r.max_steps, cap - 64, 1);
  ev = P.begin(s);
  launch_write_image(r, vs, s);
  P.end(ev, s, "write_image", n_px, n_px * 24, 0);
  const std::vector<IterLog> lg = read_iter_log(r, it, s);
  int64_t n_res = 0, n_hit = 0;
  int n_real = 0;
  for (int i = 0; i < it; i++) {
    const IterLog& l = lg[i];
    if (l.alive > 0) n_real = i + 1;
    const double bins = (double)(uint32_t)l.bins;
    // per bin: occupancy byte; per classified sample: index + density probe
    // (8 B) + 28 B stream record; ray state + entry per ray
    // march: 1 B occupancy per bin; per emitted sample its 20 B record
    // (position or froxel + pixel, t0) and the 12 B its classification reads
    // (known range, brick slot, voxel density); ray state + entry per ray
    P.work(hs[i].march, bins, bins + 32.0 * (l.n_r + l.n_c) + 84.0 * l.alive, 0);
    P.work(hs[i].net, l.n_r, l.n_r * (ALG_GATHER_B + 16.0 + 8.0), l.n_r * net_flop(m.net));
    P.work(hs[i].head, l.n_c, l.n_c * (ALG_LOOKUP_B + (m.net.split_head ? 40.0 : 20.0) + 8.0),
           l.n_c * (m.net.split_head ? ALG_VD_COLOUR_FLOP : ALG_COLOUR_FLOP));
    P.work(hs[i].comp, l.n_r + l.n_c, 12.0 * (l.n_r + l.n_c) + 132.0 * l.alive, 0);
    n_res += l.n_r;
    n_hit += l.n_c;
  }
  r.last_iters = n_real;
  r.graph_iters[1] = n_real;
  P.collect();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    const float np = (float)n_px;
    stats->samples_ppx = (float)(n_res + n_hit) / np;
    stats->resample_samples_evaluated = (float)n_res / np;
    stats->cache_samples_evaluated = (float)n_hit / np;
    stats->n_resamples = (uint64_t)n_res;
    stats->n_cache_hits = (uint64_t)n_hit;
    stats->n_subiters = n_real;
  }
}

void render_cached(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, fvc_render_stats* stats) {
  render_cached(r, Views(&rm, 1, &rgba, r.model), stats);
}

void cache_fill(Renderer& r, const fvc_raymarch_info& rm, fvc_init_stats* stats) {
  cache_fill(r, rm, stats, r.stream(), r.prof);
}

// ---------------------------------------------------------------- inspection
// The cache head's lookup alone at given froxels (one lane per froxel):
// [density row, latents 0..7] per froxel, as the head receives them.
template <bool LAT8>
__global__ __launch_bounds__(256) void k_debug_lookup(const float* __restrict__ fr, int n, CacheView cv, int di,
                                                      float* __restrict__ out9) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = i < n ? i : n - 1;
  const float4 rec = make_float4(fr[3 * j], fr[3 * j + 1], fr[3 * j + 2], 0.0f);
  _Float16 L[8], logsig;
  cache_head_inputs<LAT8>(cv, rec, di, L, logsig);
  if (i < n) {
    out9[9 * (int64_t)i] = (float)logsig;
#pragma unroll
    for (int k = 0; k < 8; k++) out9[9 * (int64_t)i + 1 + k] = (float)L[k];
  }
}

void debug_lookup(Renderer& r, int which, const float* fr_host, int64_t n, float* out_host) {
  const Model& m = *r.model;
  const CacheStore& C = r.caches[which == 0 ? r.render_idx : r.fill_idx];
  if (!C.initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
  if (n <= 0) return;
  if (n > (1 << 24)) throw Error(FVC_E_INVALID, "at most 2^24 froxels per call");
  hipStream_t s = r.stream();
  CacheView cv = cache_view(C);
  cv.interp_fn = r.settings.interpol_function;
  cv.inter_brick = r.settings.use_inter_brick_interpolation ? 1 : 0;
  DevBuf f, o;
  f.ensure(n * 3 * sizeof(float));
  o.ensure(n * 9 * sizeof(float));
  FVC_HIP(hipMemcpyAsync(f.p, fr_host, n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  const int di = m.net.first_latent_is_density ? 1 : 0;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (cv.lat8)
    hipLaunchKernelGGL(k_debug_lookup<true>, grid, dim3(256), 0, s, f.as<float>(), (int)n, cv, di, o.as<float>());
  else
    hipLaunchKernelGGL(k_debug_lookup<false>, grid, dim3(256), 0, s, f.as<float>(), (int)n, cv, di, o.as<float>());
  FVC_HIP(hipGetLastError());
  FVC_HIP(hipMemcpyAsync(out_host, o.p, n * 9 * sizeof(float), hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
}

// A cached render with the per-pixel inspection records (see k_cache_march
// and k_cache_composite): codes [pixels][k] (-1 past the ray's classified
// samples), classified [pixels], composited [pixels][2].
void debug_render_cached(Renderer& r, const fvc_raymarch_info& rm, int k, int8_t* codes_host, int32_t* classified_host,
                         int32_t* composited_host, fvc_render_stats* stats) {
  const int64_t n = (int64_t)rm.cam_info.res_x * rm.cam_info.res_y;
  if (k < 0 || (int64_t)k * n > ((int64_t)1 << 30)) throw Error(FVC_E_INVALID, "record too large");
  hipStream_t s = r.stream();
  DevBuf codes, cls, comp;
  codes.ensure(std::max<int64_t>(1, n * k));
  cls.ensure(n * sizeof(int));
  comp.ensure(n * sizeof(int2));
  FVC_HIP(hipMemsetAsync(codes.p, 0xff, std::max<int64_t>(1, n * k), s));
  FVC_HIP(hipMemsetAsync(cls.p, 0, n * sizeof(int), s));
  FVC_HIP(hipMemsetAsync(comp.p, 0, n * sizeof(int2), s));
  r.dbg_codes = codes.as<int8_t>();
  r.dbg_k = k;
  r.dbg_n = cls.as<int>();
  r.dbg_comp = comp.as<int2>();
  try {
    render_cached(r, rm, nullptr, stats);
  } catch (...) {
    r.dbg_codes = nullptr; r.dbg_n = nullptr; r.dbg_comp = nullptr; r.dbg_k = 0;
    throw;
  }
  r.dbg_codes = nullptr; r.dbg_n = nullptr; r.dbg_comp = nullptr; r.dbg_k = 0;
  if (codes_host && k > 0) FVC_HIP(hipMemcpyAsync(codes_host, codes.p, n * k, hipMemcpyDeviceToHost, s));
  if (classified_host) FVC_HIP(hipMemcpyAsync(classified_host, cls.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  if (composited_host) FVC_HIP(hipMemcpyAsync(composited_host, comp.p, n * sizeof(int2), hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- refill
// Background refill of a double-buffered renderer, the viewer's update
// thread (application.cu:168-184): the fill of the back cache runs on its
// own thread and stream (holding the write lock) while frames render from
// the front cache; a finished fill is published by a swap.
void fill_async(Renderer& r, const fvc_raymarch_info& rm) {
  if (!r.double_buffered) throw Error(FVC_E_INVALID, "an asynchronous fill needs a double-buffered renderer");
  if (r.fill_state.load() == FILL_RUNNING) throw Error(FVC_E_INVALID, "a fill is already running");
  if (r.fill_thread.joinable()) r.fill_thread.join();
  if (!r.fill_stream) FVC_HIP(hipStreamCreateWithFlags(&r.fill_stream, hipStreamNonBlocking));
  r.fill_state = FILL_RUNNING;
  r.fill_error.clear();
  try {
    r.fill_thread = std::thread([&r, rm] {
      try {
        FVC_HIP(hipSetDevice(r.model->device));
        std::scoped_lock lk(r.write_mutex);
        cache_fill(r, rm, &r.fill_stats, r.fill_stream, r.fill_prof);
        r.fill_pose = translation(rm.cam_info.cam2world);
        r.fill_state = FILL_DONE;
      } catch (const std::exception& e) {
        r.fill_error = e.what();
        r.fill_state = FILL_FAILED;
      }
    });
  } catch (...) {
    // no thread: nothing runs, and readers of the fill cache must not wait
    r.fill_state = FILL_IDLE;
    throw;
  }
}

// state of the background fill; with wait, blocks until it has finished
int fill_poll(Renderer& r, bool wait, fvc_init_stats* stats) {
  if (wait && r.fill_thread.joinable()) r.fill_thread.join();
  const int st = r.fill_state.load();
  if (st == FILL_FAILED) {
    if (r.fill_thread.joinable()) r.fill_thread.join();
    r.fill_state = FILL_IDLE;
    throw Error(FVC_E_HIP, "background fill failed: " + r.fill_error);
  }
  if (st == FILL_DONE && stats) *stats = r.fill_stats;
  return st;
}

// publish a finished background fill (swap under both locks)
bool fill_publish(Renderer& r) {
  if (r.fill_state.load() != FILL_DONE) return false;
  if (r.fill_thread.joinable()) r.fill_thread.join();
  std::scoped_lock lk(r.read_mutex, r.write_mutex);
  std::swap(r.render_idx, r.fill_idx);
  r.front_pose = r.fill_pose;
  r.have_front_pose = true;
  r.fill_state = FILL_IDLE;
  return true;
}

// Render from the cache and apply the refill policy of the viewer
// (application.cu:203-213): refill when the frame's cache hit ratio drops
// below min_hit_ratio or the camera is farther than max_distance from the
// pose of the front cache.  The first frame fills synchronously.
void render_cached_auto(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, fvc_render_stats* stats,
                        int* refill_started) {
  if (refill_started) *refill_started = 0;
  fill_publish(r);
  if (!r.caches[r.render_idx].initialized) {
    {
      std::scoped_lock lk(r.write_mutex);
      cache_fill(r, rm, nullptr);
    }
    std::scoped_lock lk(r.read_mutex, r.write_mutex);
    if (r.double_buffered) std::swap(r.render_idx, r.fill_idx);
    r.front_pose = translation(rm.cam_info.cam2world);
    r.have_front_pose = true;
  }
  fvc_render_stats st{};
  {
    std::scoped_lock lk(r.read_mutex);
    render_cached(r, rm, rgba, &st);
  }
  if (stats) *stats = st;
  const fvc_refill_policy& pol = r.policy;
  if (!pol.enabled) return;
  const float hit = st.samples_ppx > 0.0f ? st.cache_samples_evaluated / st.samples_ppx : 1.0f;
  const V3 d = translation(rm.cam_info.cam2world) - r.front_pose;
  const bool moved = r.have_front_pose && std::sqrt(dot3(d, d)) > pol.max_distance;
  if ((hit < pol.min_hit_ratio || moved) && r.fill_state.load() == FILL_IDLE) {
    if (r.double_buffered) {
      fill_async(r, rm);
    } else {
      std::scoped_lock lk(r.read_mutex, r.write_mutex);
      cache_fill(r, rm, nullptr);
      r.front_pose = translation(rm.cam_info.cam2world);
    }
    if (refill_started) *refill_started = 1;
  }
}

}  // namespace fvc
