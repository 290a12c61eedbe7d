// Domain set-up on the device (backend/crates/orchestrator/src/domain.rs:
// 35-402) and analytic GridSDF node values.  Included by ubp_engine.hip.
//
// The set-up visits the lattice sites (i + 0.5) spacing, i outer and k inner
// (domain.rs:111-157), in blocks of 1024 sites: a count pass gives every
// block its number of fluid sites (inside the domain box, SDF >= 0) and of
// geometry sites (0 < SDF < spacing / 2); the host scans the block counts;
// a write pass recomputes the flags, ranks each site inside its block and
// writes it at its place in the reference's order.  Wall layers are a closed
// form (domain.rs:162-348), one thread per particle.  The SDF is sampled as
// the reference does, f32::MAX outside the node grid (geometry.rs:99-175).

namespace ubp {

struct DomSdf {
  float origin[3];
  float cell;
  uint32_t dims[3];
  const float* data;
  uint64_t count;
};

struct DomDev {
  float dmin[3], dmax[3];
  float spacing, temperature, center_y, water_mass, air_mass, bmass, thr;
  int fluid_type;
  uint32_t cnt[3];
  DomSdf S;
};

// `f.floor() as i32`: saturating, NaN -> 0
__device__ __forceinline__ int32_t dom_sat_i32(float v) {
  if (v != v) return 0;
  if (v <= -2147483648.0f) return INT32_MIN;
  if (v >= 2147483648.0f) return INT32_MAX;
  return int32_t(v);
}
__device__ __forceinline__ float dom_node(const DomSdf& S, uint32_t i, uint32_t j, uint32_t k) {
  const uint32_t idx = i * S.dims[1] * S.dims[2] + j * S.dims[2] + k;  // u32, wrapping
  return uint64_t(idx) < S.count ? S.data[idx] : FLT_MAX;
}
// GridSDF::sample (geometry.rs:99-150)
__device__ float dom_sample(const DomSdf& S, float px, float py, float pz) {
  const float gx = (px - S.origin[0]) / S.cell;
  const float gy = (py - S.origin[1]) / S.cell;
  const float gz = (pz - S.origin[2]) / S.cell;
  const int32_t i0 = dom_sat_i32(floorf(gx)), j0 = dom_sat_i32(floorf(gy)), k0 = dom_sat_i32(floorf(gz));
  if (i0 < 0 || j0 < 0 || k0 < 0) return FLT_MAX;
  const int32_t i1 = int32_t(uint32_t(i0) + 1u), j1 = int32_t(uint32_t(j0) + 1u), k1 = int32_t(uint32_t(k0) + 1u);
  if (i1 >= int32_t(S.dims[0]) || j1 >= int32_t(S.dims[1]) || k1 >= int32_t(S.dims[2])) return FLT_MAX;
  const float fx = gx - float(i0), fy = gy - float(j0), fz = gz - float(k0);
  const uint32_t a0 = uint32_t(i0), b0 = uint32_t(j0), c0_ = uint32_t(k0);
  const uint32_t a1 = uint32_t(i1), b1 = uint32_t(j1), c1_ = uint32_t(k1);
  const float c000 = dom_node(S, a0, b0, c0_), c001 = dom_node(S, a0, b0, c1_);
  const float c010 = dom_node(S, a0, b1, c0_), c011 = dom_node(S, a0, b1, c1_);
  const float c100 = dom_node(S, a1, b0, c0_), c101 = dom_node(S, a1, b0, c1_);
  const float c110 = dom_node(S, a1, b1, c0_), c111 = dom_node(S, a1, b1, c1_);
  const float c00 = c000 * (1.0f - fz) + c001 * fz;
  const float c01 = c010 * (1.0f - fz) + c011 * fz;
  const float c10 = c100 * (1.0f - fz) + c101 * fz;
  const float c11 = c110 * (1.0f - fz) + c111 * fz;
  const float c0 = c00 * (1.0f - fy) + c01 * fy;
  const float c1 = c10 * (1.0f - fy) + c11 * fy;
  return c0 * (1.0f - fx) + c1 * fx;
}
// GridSDF::gradient (geometry.rs:153-175)
__device__ void dom_gradient(const DomSdf& S, float px, float py, float pz, float g[3]) {
  const float h = S.cell * 0.5f;
  const float dx = (dom_sample(S, px + h, py, pz) - dom_sample(S, px - h, py, pz)) / (2.0f * h);
  const float dy = (dom_sample(S, px, py + h, pz) - dom_sample(S, px, py - h, pz)) / (2.0f * h);
  const float dz = (dom_sample(S, px, py, pz + h) - dom_sample(S, px, py, pz - h)) / (2.0f * h);
  const float mag = sqrtf(dx * dx + dy * dy + dz * dz);
  if (mag > 1e-8f) {
    g[0] = dx / mag; g[1] = dy / mag; g[2] = dz / mag;
  } else {
    g[0] = g[1] = g[2] = 0.0f;
  }
}

// site -> position and flags (fluid: inside the domain box and SDF >= 0;
// geometry: 0 < SDF < spacing / 2)
__device__ __forceinline__ void dom_site(const DomDev& D, uint32_t site, float& x, float& y, float& z, bool& fl,
                                         bool& ge) {
  const uint32_t plane = D.cnt[1] * D.cnt[2];
  const uint32_t i = site / plane, rem = site - i * plane;
  const uint32_t j = rem / D.cnt[2], k = rem - j * D.cnt[2];
  x = D.dmin[0] + (float(i) + 0.5f) * D.spacing;
  y = D.dmin[1] + (float(j) + 0.5f) * D.spacing;
  z = D.dmin[2] + (float(k) + 0.5f) * D.spacing;
  const float d = dom_sample(D.S, x, y, z);
  const bool inside = !(x < D.dmin[0] || x > D.dmax[0] || y < D.dmin[1] || y > D.dmax[1] || z < D.dmin[2] ||
                        z > D.dmax[2]);
  fl = inside && !(d < 0.0f);
  ge = d > 0.0f && d < D.thr;
}

#define UBP_DOM_SITES 1024u  // sites per block: 4 per thread of 256

__global__ __launch_bounds__(256) void k_dom_count(DomDev D, uint32_t nsites, uint32_t* __restrict__ blk_f,
                                                   uint32_t* __restrict__ blk_g) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * UBP_DOM_SITES + threadIdx.x * 4u;
  uint32_t nf = 0, ng = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (base + q < nsites) {
      float x, y, z;
      bool fl, ge;
      dom_site(D, base + q, x, y, z, fl, ge);
      nf += fl ? 1u : 0u;
      ng += ge ? 1u : 0u;
    }
  }
  uint32_t tf = 0, tg = 0;
  (void)block_excl_scan256(nf, sh, tf);
  (void)block_excl_scan256(ng, sh, tg);
  if (threadIdx.x == 0) {
    blk_f[blockIdx.x] = tf;
    blk_g[blockIdx.x] = tg;
  }
}

struct DomOut {
  float *fx, *fy, *fz, *fm, *frho, *ftemp;
  uint8_t* ftype;
  float *bx, *by, *bz, *bm, *bnx, *bny, *bnz;
};

__global__ __launch_bounds__(256) void k_dom_write(DomDev D, uint32_t nsites, const uint32_t* __restrict__ off_f,
                                                   const uint32_t* __restrict__ off_g, uint32_t nwall, DomOut O) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * UBP_DOM_SITES + threadIdx.x * 4u;
  float px[4], py[4], pz[4];
  bool fl[4], ge[4];
  uint32_t nf = 0, ng = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    fl[q] = ge[q] = false;
    px[q] = py[q] = pz[q] = 0.0f;
    if (base + q < nsites) dom_site(D, base + q, px[q], py[q], pz[q], fl[q], ge[q]);
    nf += fl[q] ? 1u : 0u;
    ng += ge[q] ? 1u : 0u;
  }
  uint32_t tf = 0, tg = 0;
  uint32_t rf = off_f[blockIdx.x] + block_excl_scan256(nf, sh, tf);
  uint32_t rg = off_g[blockIdx.x] + block_excl_scan256(ng, sh, tg);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (fl[q]) {
      const bool water = D.fluid_type == 0 || (D.fluid_type == 2 && py[q] < D.center_y);
      O.fx[rf] = px[q]; O.fy[rf] = py[q]; O.fz[rf] = pz[q];
      O.fm[rf] = water ? D.water_mass : D.air_mass;
      O.frho[rf] = water ? 1000.0f : 1.2f;
      O.ftemp[rf] = D.temperature;
      O.ftype[rf] = water ? 0 : 1;
      ++rf;
    }
    if (ge[q]) {
      float g[3];
      dom_gradient(D.S, px[q], py[q], pz[q], g);
      for (int layer = 0; layer < 3; ++layer) {
        const size_t b = size_t(nwall) + 3u * size_t(rg) + size_t(layer);
        const float off = (float(layer) + 0.5f) * D.spacing;
        O.bx[b] = px[q] - g[0] * off;
        O.by[b] = py[q] - g[1] * off;
        O.bz[b] = pz[q] - g[2] * off;
        O.bm[b] = D.bmass;
        O.bnx[b] = g[0]; O.bny[b] = g[1]; O.bnz[b] = g[2];
      }
      ++rg;
    }
  }
}

// wall particle w: face (enabled faces in order x-, x+, y-, y+, z-, z+, each
// 3 * cnt[u] * cnt[v] particles: layer, then the u index, then the v index)
__global__ UBP_LB void k_dom_walls(DomDev D, uint32_t walls, uint32_t nwall, DomOut O) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwall) return;
  uint32_t rem = w;
  int f = 0;
  for (; f < 6; ++f) {
    if (!(walls & (1u << f))) continue;
    const int a = f / 2, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
    const uint32_t nface = 3u * D.cnt[u] * D.cnt[v];
    if (rem < nface) break;
    rem -= nface;
  }
  const int a = f / 2, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
  const bool hi = (f & 1) != 0;
  const uint32_t per_layer = D.cnt[u] * D.cnt[v];
  const uint32_t layer = rem / per_layer, pq = rem - layer * per_layer;
  const uint32_t p = pq / D.cnt[v], q = pq - p * D.cnt[v];
  float c[3], n[3] = {0.0f, 0.0f, 0.0f};
  c[u] = D.dmin[u] + (float(p) + 0.5f) * D.spacing;
  c[v] = D.dmin[v] + (float(q) + 0.5f) * D.spacing;
  const float off = (float(layer) + 0.5f) * D.spacing;
  c[a] = hi ? D.dmax[a] + off : D.dmin[a] - off;
  n[a] = hi ? -1.0f : 1.0f;
  O.bx[w] = c[0]; O.by[w] = c[1]; O.bz[w] = c[2];
  O.bm[w] = D.bmass;
  O.bnx[w] = n[0]; O.bny[w] = n[1]; O.bnz[w] = n[2];
}

// Analytic node values in FP64, rounded once: node (i, j, k) at
// origin + cell (i, j, k), z fastest.  kind 0: distance inside the box
// [p0..p2, p3..p5], -(|max(q, 0)| + min(max_a q_a, 0)) with
// q = max(lo - x, x - hi); kind 1: |x - c| - r with c = p0..p2, r = p3, and
// with `combine` the minimum with the value already stored.
__global__ UBP_LB void k_sdf_analytic(int kind, const double* __restrict__ prm, double o0, double o1, double o2,
                                      double cell, uint32_t d0, uint32_t d1, uint32_t d2, int combine,
                                      float* __restrict__ out) {
  const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t total = uint64_t(d0) * d1 * d2;
  if (t >= total) return;
  const uint32_t i = uint32_t(t / (uint64_t(d1) * d2));
  const uint32_t rem = uint32_t(t - uint64_t(i) * d1 * d2);
  const uint32_t j = rem / d2, k = rem - j * d2;
  const double x = o0 + cell * double(i), y = o1 + cell * double(j), z = o2 + cell * double(k);
  double d;
  if (kind == 0) {
    const double qx = fmax(prm[0] - x, x - prm[3]);
    const double qy = fmax(prm[1] - y, y - prm[4]);
    const double qz = fmax(prm[2] - z, z - prm[5]);
    const double mx = fmax(qx, 0.0), my = fmax(qy, 0.0), mz = fmax(qz, 0.0);
    const double outside = sqrt((mx * mx + my * my) + mz * mz);
    const double inside = fmin(fmax(fmax(qx, qy), qz), 0.0);
    d = -(outside + inside);
  } else {
    const double ex = x - prm[0], ey = y - prm[1], ez = z - prm[2];
    d = sqrt((ex * ex + ey * ey) + ez * ez) - prm[3];
  }
  const float v = float(d);
  out[t] = (combine && !(v < out[t])) ? out[t] : v;
}

}  // namespace ubp
