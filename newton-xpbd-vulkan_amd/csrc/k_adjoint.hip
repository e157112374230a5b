// k_adjoint.hip -- reverse-mode kernels of one substep (one iteration) for
// the differentiable rollout.  No reference counterpart: the reference has no
// gradient path.  The forward substep is
//   v1 = clamp(v0 + g dt),  x1 = x0 + v1 dt                 (integrate)
//   D  = relax * w_i * ( sum_shape g_s / (w_i + w_b) + sum_pairs g_ij / (w_i + w_j) )
//   x2 = x1 + D,  v2 = (x2 - x0) / dt  (clamped)             (apply)
// with the contact set frozen from the forward pass.  Per particle-pair term
//   g(rd, vrel) = t * max(mu err, -|vt| dt) - n err,  n = rd/|rd|,
//   err = |rd| - r_i - r_j, vt = vrel - (n.vrel) n, t = vt/|vt|
// (solve_particle_particle_contacts.comp:170-205) and per shape contact the
// same form with c = n.(x1 - b) - r and a constant normal
// (solve_particle_shape_contacts.comp:143-180).  g is odd in (rd, vrel), so
// the pair (i, j) enters particle k's cotangent through one Jacobian with the
// combined cotangent s_kj a_k - s_jk a_j: the particle-pair adjoint is a
// gather over the same neighbour boxes as the forward solve, without atomics.
#include "xpbd_kernels.h"

namespace xpbd {

__device__ __forceinline__ f3 perp3(f3 u, f3 n) { return u - n * dot3(n, u); }

// VJP of the pair term at (rd, vrel) with cotangent C; returns g.
__device__ __forceinline__ f3 pair_vjp(f3 rd, f3 vrel, float d, float err, float kmu, float h, f3 C, f3& g_rd, f3& g_vr) {
    const float invd = 1.0f / d;
    const f3 n = rd * invd;
    const float vn = dot3(n, vrel);
    const f3 vt = vrel - n * vn;
    const float L = len3(vt);
    const float la = kmu * err, lb = -L * h;
    const f3 pc = perp3(C, n);
    const float nc = dot3(n, C);
    if (!(la < lb)) {
        const f3 t = L > 0.0f ? vt * (1.0f / L) : f3{0.f, 0.f, 0.f};
        g_rd = n * (kmu * dot3(t, C) - nc) - pc * (err * invd);
        g_vr = {0.f, 0.f, 0.f};
        if (L > 0.0f) {
            const f3 u = (C - t * dot3(t, C)) * (kmu * err / L);
            const f3 pu = perp3(u, n);
            g_vr = pu;
            g_rd = g_rd - (vt * dot3(n, u) + pu * vn) * invd;
        }
        return t * la - n * err;
    }
    g_vr = pc * (-h);
    g_rd = (vt * nc + pc * vn) * (h * invd) - n * nc - pc * (err * invd);
    return vt * (-h) - n * err;
}

// VJP of the shape-contact term; returns g.
__device__ __forceinline__ f3 shape_vjp(float c, f3 v, f3 n, float mu, float h, f3 C, f3& g_x, f3& g_v) {
    const float vn = dot3(n, v);
    const f3 vt = v - n * vn;
    const float L = len3(vt);
    const float la = mu * c, lb = -L * h;
    if (!(la < lb)) {
        const f3 t = L > 0.0f ? vt * (1.0f / L) : f3{0.f, 0.f, 0.f};
        g_x = n * (mu * dot3(t, C) - dot3(n, C));
        g_v = L > 0.0f ? perp3((C - t * dot3(t, C)) * (mu * c / L), n) : f3{0.f, 0.f, 0.f};
        return t * la - n * c;
    }
    g_x = n * (-dot3(n, C));
    g_v = perp3(C, n) * (-h);
    return vt * (-h) - n * c;
}

// Adds the block's sum of v into *slot (one slot per block: no atomics, and
// the same order every launch).
__device__ __forceinline__ void block_accumulate(double v, double* slot) {
    __shared__ double s_red[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (__lane_id() == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < int(blockDim.x >> 6); k++) t += s_red[k];
        *slot += t;
    }
}

// Apply adjoint (apply_particle_deltas.comp:44-57) per sorted slot t: the
// cotangent a of the unclamped x1 + D and the apply part of the x0 cotangent.
// A substep moves the state from the previous order (slot vals[t]: x0, v0 and
// their cotangents) to the sorted order (slot t: the end state, its
// cotangents, deltas and contacts).
__global__ __launch_bounds__(256) void k_adj_prep(AdjArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const int i = int(a.vals[t]);
    if ((a.Fs[t] & 1) == 0) {
        a.As[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const f3 x1 = xyz(a.Xs[t]), x0 = xyz(a.P0[i]);
    const f3 dl = {a.delta[3 * t], a.delta[3 * t + 1], a.delta[3 * t + 2]};
    const f3 xn = x1 + dl;
    const f3 u = markstein_div(xn - x0, a.dt);
    const float m = len3(u);
    const f3 lx = xyz(a.lx_in[t]), lv = xyz(a.lv_in[t]);
    f3 av, lxa;
    if (m > a.vmax) {
        const f3 uh = u * (1.0f / m);
        const f3 lt = lv + lx * a.dt;
        const f3 lu = (lt - uh * dot3(uh, lt)) * (a.vmax / m);
        av = lu * (1.0f / a.dt);
        lxa = lx - av;
    } else {
        av = lx + lv * (1.0f / a.dt);
        lxa = lv * (-1.0f / a.dt);
    }
    a.As[t] = make_float4(av.x, av.y, av.z, 0.f);
    a.lx_out[i] = make_float4(lxa.x, lxa.y, lxa.z, 0.f);
}

// Shape-contact adjoint, one thread per contact (the forward's skip rules);
// cotangents of x1 and v1 accumulate into shx / shv.
__global__ __launch_bounds__(256) void k_adj_shape(AdjShapeArgs a) {
    const int count = min(a.cmax, *a.cc);
    double racc = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const int s = a.cshape[t];
        const int body = a.body.shape_body[s];
        const int pi = a.cpart[t];
        const int pf = a.flags[pi];
        if ((pf & 1) == 0) continue;
        if (pf & 2) {
            if (body < 0) continue;
            if (a.body.flags[body] & 4) continue;
            if (a.body.invm[body] == 0.0f) continue;
        }
        const float4 P = a.Xs[pi], Vv = a.Vs[pi];   // contacts name sorted slots
        const f3 px = xyz(P), pv = xyz(Vv);
        const BodyFrame F = load_body(a.body, body);
        const f3 bx = F.pos + quat_rotate_sc(F.rot, xyz(a.cbpos[t]));
        const f3 rr = bx - (F.pos + quat_rotate_sc(F.rot, F.com));
        const f3 n = xyz(a.cnorm[t]);
        const f3 d = px - bx;
        const float c = fma_(n.z, d.z, fma_(n.x, d.x, n.y * d.y)) - P.w;
        if (c > a.ka) continue;
        const float w1 = Vv.w;
        float w2 = 0.0f;
        if (body >= 0) {
            const f3 ang = cross3(rr, n);
            const f3 ra = quat_rotate_inv_sc(F.rot, ang);
            const float* I = a.body.invI + 9 * body;
            const f3 Ira = {I[0] * ra.x + I[1] * ra.y + I[2] * ra.z, I[3] * ra.x + I[4] * ra.y + I[5] * ra.z,
                            I[6] * ra.x + I[7] * ra.y + I[8] * ra.z};
            w2 = a.body.invm[body] + dot3(ra, Ira);
        }
        const float denom = w1 + w2;
        if (denom == 0.0f) continue;
        const float mu = 0.5f * (a.pmu + a.body.mat_mu[s]);
        const f3 bv = F.v + cross3(F.w, rr) + quat_rotate_sc(F.rot, xyz(a.cbvel[t]));
        const float sc = w1 / denom;
        const f3 av = xyz(a.As[pi]);
        f3 g_x, g_v;
        const f3 g = shape_vjp(c, pv - bv, n, mu, a.dt, av * (sc * a.relax), g_x, g_v);
        atomicAdd(&a.shx[3 * pi + 0], g_x.x); atomicAdd(&a.shx[3 * pi + 1], g_x.y); atomicAdd(&a.shx[3 * pi + 2], g_x.z);
        atomicAdd(&a.shv[3 * pi + 0], g_v.x); atomicAdd(&a.shv[3 * pi + 1], g_v.y); atomicAdd(&a.shv[3 * pi + 2], g_v.z);
        racc += double(dot3(av, g) * sc);
    }
    block_accumulate(racc, a.relax_part + blockIdx.x);
}

// Particle-pair adjoint gathered per sorted slot, then the integrate adjoint
// (integrate_particles.comp:44-57); writes the substep-start cotangents.
constexpr int kSurvAdj = 16;   // recorded candidates per lane before a drain

template <bool BATCHED>
__global__ __launch_bounds__(256) void k_adj_pp(AdjArgs a) {
    __shared__ uint32_t s_surv[kSurvAdj][256];
    const int t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double racc = 0.0;
    if (t < a.n) {
        const int k = int(a.vals[t]);
        const int fk = a.Fs[t];
        if ((fk & 1) == 0) {
            a.lx_out[k] = a.lx_in[t];
            a.lv_out[k] = a.lv_in[t];
        } else {
            const float4 Pk = a.Xs[t], Vk = a.Vs[t];
            const f3 x = xyz(Pk), v = xyz(Vk), ak = xyz(a.As[t]);
            const float rk = Pk.w, wk = Vk.w;
            const int proxy_k = fk & 2;
            const float R = rk + a.max_r + a.coh;
            const float inv = a.inv_cell;
            const int xs = (int)floorf((x.x - R) * inv), ys = (int)floorf((x.y - R) * inv), zs = (int)floorf((x.z - R) * inv);
            const int xe = min((int)floorf((x.x + R) * inv), xs + kGridDim - 1);
            const int ye = min((int)floorf((x.y + R) * inv), ys + kGridDim - 1);
            const int ze = min((int)floorf((x.z + R) * inv), zs + kGridDim - 1);
            const int2* range = reinterpret_cast<const int2*>(a.range);
            const uint32_t base = BATCHED ? (a.keys[t] & ~((1u << kCellBits) - 1u)) : 0u;
            f3 gx = {0.f, 0.f, 0.f}, gv = {0.f, 0.f, 0.f};
            float rsum = 0.0f;
            // the forward's cell skipping (every active pair, from either side,
            // lies within r_k + r_max + |coh|); the rollout runs one iteration,
            // so the positions are those the cells were computed from
            const float ux = x.x * inv, uy = x.y * inv, uz = x.z * inv;
            const float lu = (rk + a.rmax + fabsf(a.coh)) * inv;
            const float lu2 = lu * lu * 1.0001f + 1e-6f;
            auto gap2 = [&](float u, int c) -> float {
                if (c + kGridOrigin < 0) return 0.0f;
                const float lo = float(c), hi = float(c) + 1.0f;
                float g = lo > u ? lo - u : (u >= hi ? u - hi : 0.0f);
                g -= 1e-5f * (fabsf(u) + 1.0f);
                return g > 0.0f ? g * g : 0.0f;
            };
            // two phases as in the forward solve: a cheap-reject scan records
            // candidate slots per lane, the reverse projection runs over them
            const float lim0 = rk + fabsf(a.coh);
            int cnt = 0;
            auto drain = [&]() {
                for (int q = 0; q < cnt; q++) {
                    const int m = int(s_surv[q][threadIdx.x]);
                    const int nf = a.Fs[m];
                    if ((nf & 1) == 0) continue;
                    const float4 Pj = a.Xs[m], Vj = a.Vs[m];
                    const f3 rd = x - xyz(Pj);
                    const float ax = fmaxf(fmaxf(fabsf(rd.x), fabsf(rd.y)), fabsf(rd.z));
                    const float rj = Pj.w, wj = Vj.w;
                    const float Rj = rj + a.max_r + a.coh;
                    if (ax > R && ax > Rj) continue;
                    const float d = exact_length(rd);
                    if (!(d > 0.0f)) continue;
                    // particle k's solve and particle j's solve, each with its own skip rules
                    const float err_kj = d - rk - rj, err_jk = d - rj - rk;
                    const bool kj = !(proxy_k != 0 && ((nf & 2) != 0 || wj <= 0.0f)) && !(ax > R) &&
                                    !(err_kj > a.coh) && (wk + wj) > 0.0f;
                    const bool jk = !((nf & 2) != 0 && (proxy_k != 0 || wk <= 0.0f)) && !(ax > Rj) &&
                                    !(err_jk > a.coh) && (wj + wk) > 0.0f;
                    if (!kj && !jk) continue;
                    const float den = wk + wj;
                    f3 C = {0.f, 0.f, 0.f};
                    if (kj) C = ak * (wk * a.relax / den);
                    if (jk) C = C - xyz(a.As[m]) * (wj * a.relax / den);
                    f3 g_rd, g_vr;
                    const f3 g = pair_vjp(rd, v - xyz(Vj), d, err_kj, a.mu, a.dt, C, g_rd, g_vr);
                    gx = gx + g_rd;
                    gv = gv + g_vr;
                    if (kj) rsum += dot3(ak, g) * (wk / den);
                }
                cnt = 0;
            };
            for (int cz = zs; cz <= ze; cz++)
                for (int cy = ys; cy <= ye; cy++)
                    for (int cx = xs; cx <= xe; cx++) {
                        if (gap2(ux, cx) + gap2(uy, cy) + gap2(uz, cz) > lu2) continue;
                        const uint32_t tb = BATCHED ? base : (a.planes ? wrap_plane(cx, cy, cz, a.planes) << kCellBits : 0u);
                        const uint32_t cell = tb + uint32_t(cell_index(cx, cy, cz));
                        const int2 rg = BATCHED ? row_cell(a.rowtab, cell) : range[size_t(cell)];
                        for (int m = rg.x; m < rg.y; m++) {
                            const float4 Pj = a.Xs[m];
                            const float dx = x.x - Pj.x, dy = x.y - Pj.y, dz = x.z - Pj.z;
                            const float lim = lim0 + Pj.w;   // as the forward's cheap reject
                            if (m == t || (dx * dx + dy * dy) + dz * dz > lim * lim * 1.0001f + 1e-30f) continue;
                            s_surv[cnt][threadIdx.x] = uint32_t(m);
                            if (++cnt == kSurvAdj) drain();
                        }
                    }
            drain();
            racc = double(rsum);
            const f3 sx = {a.shx[3 * t], a.shx[3 * t + 1], a.shx[3 * t + 2]};
            const f3 sv = {a.shv[3 * t], a.shv[3 * t + 1], a.shv[3 * t + 2]};
            a.shx[3 * t] = 0.f; a.shx[3 * t + 1] = 0.f; a.shx[3 * t + 2] = 0.f;
            a.shv[3 * t] = 0.f; a.shv[3 * t + 1] = 0.f; a.shv[3 * t + 2] = 0.f;
            const f3 lx1 = ak + sx + gx;
            const f3 lv1 = sv + gv;
            const f3 lx0 = xyz(a.lx_out[k]) + lx1;
            const f3 lvv = lv1 + lx1 * a.dt;
            // integrate: recompute the clamp decision exactly as the forward
            const float4 V0 = a.V0[k];
            const float gs = V0.w > 0.0f ? 1.0f : 0.0f;
            const float fx = 0.0f * V0.w, fy = 0.0f * V0.w, fz = 0.0f * V0.w;
            const f3 v1 = {V0.x + (fx + a.gx * gs) * a.dt, V0.y + (fy + a.gy * gs) * a.dt, V0.z + (fz + a.gz * gs) * a.dt};
            const float mag = len3(v1);
            f3 lv0 = lvv;
            if (mag > a.vmax) {
                const f3 uh = v1 * (1.0f / mag);
                lv0 = (lvv - uh * dot3(uh, lvv)) * (a.vmax / mag);
            }
            a.lx_out[k] = make_float4(lx0.x, lx0.y, lx0.z, 0.f);
            a.lv_out[k] = make_float4(lv0.x, lv0.y, lv0.z, 0.f);
        }
    }
    block_accumulate(racc, a.relax_part + blockIdx.x);
}

// float3 <-> float4 state transfers of the rollout boundary (w kept).
__global__ __launch_bounds__(256) void k_xyz_in(float4* __restrict__ dst, const float* __restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 o = dst[i];
    dst[i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], o.w);
}
__global__ __launch_bounds__(256) void k_xyz_out(float* __restrict__ dst, const float4* __restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 o = src[i];
    dst[3 * i] = o.x; dst[3 * i + 1] = o.y; dst[3 * i + 2] = o.z;
}
__global__ __launch_bounds__(256) void k_fill_f4(float4* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

namespace {
inline unsigned blocks256(int n) { return unsigned((n + 255) / 256); }
}

void launch_adj_prep(const AdjArgs& a, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_adj_prep, dim3(blocks256(a.n)), dim3(256), 0, s, a);
}
void launch_adj_shape(const AdjShapeArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_shape, dim3(grid), dim3(256), 0, s, a);
}
void launch_adj_pp(const AdjArgs& a, bool batched, hipStream_t s) {
    if (a.n <= 0) return;
    if (batched) hipLaunchKernelGGL(k_adj_pp<true>, dim3(blocks256(a.n)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_adj_pp<false>, dim3(blocks256(a.n)), dim3(256), 0, s, a);
}
void launch_xyz_in(float4* dst, const float* src, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_xyz_in, dim3(blocks256(n)), dim3(256), 0, s, dst, src, n);
}
void launch_xyz_out(float* dst, const float4* src, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_xyz_out, dim3(blocks256(n)), dim3(256), 0, s, dst, src, n);
}
void launch_fill_f4(float4* dst, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f4, dim3(blocks256(n)), dim3(256), 0, s, dst, n);
}

}  // namespace xpbd
