// rollout.cpp -- differentiable rollout of the particle step (no reference
// counterpart): a forward that keeps state checkpoints and a reverse pass
// that recomputes each checkpoint segment with the plain substep, records
// what the reverse kernels read, and runs them substep by substep backwards.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "devbuf.h"

namespace xpbd {

void Solver::free_tape() {
    Tape& t = tape_;
    auto fv = [](auto& v) { for (auto& b : v) dfree(b); v.clear(); };
    fv(t.ckP); fv(t.ckV); fv(t.ckF); fv(t.ckI); fv(t.P0); fv(t.V0); fv(t.Xs); fv(t.Vs); fv(t.cbpos); fv(t.cbvel); fv(t.cnorm);
    fv(t.keys); fv(t.vals); fv(t.Fs); fv(t.cpart); fv(t.cshape); fv(t.delta);
    dfree(t.counts);
    for (int k = 0; k < 2; k++) { dfree(t.lx[k]); dfree(t.lv[k]); }
    dfree(t.As); dfree(t.finP); dfree(t.finV); dfree(t.finF); dfree(t.finI); dfree(t.shx); dfree(t.shv);
    dfree(t.relax_part);
    t.valid = false;
}

// Per-substep records of one segment plus the reverse buffers.
void Solver::ensure_tape() {
    Tape& t = tape_;
    const size_t k = size_t(t.every);
    auto grow = [&](auto& v, size_t n) {
        while (v.size() < k) { v.emplace_back(); dalloc(v.back(), n); }
    };
    grow(t.P0, n_); grow(t.V0, n_); grow(t.Xs, n_); grow(t.Vs, n_);
    grow(t.cbpos, cmax_); grow(t.cbvel, cmax_); grow(t.cnorm, cmax_);
    grow(t.keys, n_); grow(t.vals, n_); grow(t.Fs, n_); grow(t.cpart, cmax_); grow(t.cshape, cmax_);
    grow(t.delta, 3 * n_);
    if (t.counts.n < k) dalloc(t.counts, k);
    if (t.As.n != n_) {
        for (int q = 0; q < 2; q++) { dalloc(t.lx[q], n_); dalloc(t.lv[q], n_); }
        dalloc(t.As, n_); dalloc(t.finP, n_); dalloc(t.finV, n_); dalloc(t.finF, n_); dalloc(t.finI, n_);
        dalloc(t.shx, 3 * n_); dalloc(t.shv, 3 * n_);
    }
}

// Exchanges the live stage buffers with segment record j: a substep run in
// between writes its sorted arrays, contacts and deltas into the record (its
// sorted flags are copied after it).
void Solver::swap_tape(int j) {
    Tape& t = tape_;
    std::swap(keysA_, t.keys[j]); std::swap(valsA_, t.vals[j]);
    std::swap(Xs_[0], t.Xs[j]); std::swap(Vs_[0], t.Vs[j]);
    std::swap(delta_, t.delta[j]);
    std::swap(cpart_, t.cpart[j]); std::swap(cshape_, t.cshape[j]);
    std::swap(cbpos_, t.cbpos[j]); std::swap(cbvel_, t.cbvel[j]); std::swap(cnorm_, t.cnorm[j]);
}

void Solver::rollout_forward(const float* x0, const float* v0, float relax, float sdt, int substeps, int every,
                             float* x_out, float* v_out) {
    require_init();
    if (!finalized_) finalize_meshes();
    if (!boxes_.empty()) throw std::runtime_error("rollout: dynamic bodies are not differentiated");
    if (cfg_.enable_restitution) throw std::runtime_error("rollout: restitution is not differentiated");
    if (substeps <= 0) throw std::runtime_error("rollout: substeps must be positive");
    if (!(sdt > 0.0f)) throw std::runtime_error("rollout: substep length must be positive");
    every = std::max(1, std::min(every, substeps));
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    const int n = int(n_);
    // the state is in slot order: slot t holds particle ids[t]
    if (x0) launch_gather_f3(P_[s_].p, x0, ids_cur(), n, stream_);
    if (v0) launch_gather_f3(V_[s_].p, v0, ids_cur(), n, stream_);
    Tape& t = tape_;
    t.valid = false;
    const int G = (substeps + every - 1) / every;
    if (t.every != every) {   // segment records are sized by the interval
        free_tape();
    }
    t.sdt = sdt; t.relax = relax; t.substeps = substeps; t.every = every;
    while (int(t.ckP.size()) < G) {
        t.ckP.emplace_back(); dalloc(t.ckP.back(), n_);
        t.ckV.emplace_back(); dalloc(t.ckV.back(), n_);
        t.ckF.emplace_back(); dalloc(t.ckF.back(), n_);
        t.ckI.emplace_back(); dalloc(t.ckI.back(), n_);
    }
    const float keep = relax_;
    relax_ = relax;
    try {
        launch_frame_begin(cc_.p, mcnt_.p, int(mcnt_.n), stream_);
        for (int sub = 0; sub < substeps; sub++) {
            if (sub % every == 0) {
                const int c = sub / every;
                HIPCHECK(hipMemcpyAsync(t.ckP[c].p, P_[s_].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
                HIPCHECK(hipMemcpyAsync(t.ckV[c].p, V_[s_].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
                HIPCHECK(hipMemcpyAsync(t.ckF[c].p, flags_cur(), n_ * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
                HIPCHECK(hipMemcpyAsync(t.ckI[c].p, ids_cur(), n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
            }
            substep(sdt, 1, true, sub & 1, true);
        }
    } catch (...) {
        relax_ = keep;
        throw;
    }
    relax_ = keep;
    contact_slot_ = rigid_slot_ = (substeps - 1) & 1;
    if (x_out) launch_scatter_f3(x_out, P_[s_].p, ids_cur(), n, stream_);
    if (v_out) launch_scatter_f3(v_out, V_[s_].p, ids_cur(), n, stream_);
    read_count(contact_slot_);
    t.valid = true;
}

void Solver::rollout_backward(const float* gx, const float* gv, float* gx0, float* gv0, double* grelax) {
    require_init();
    Tape& t = tape_;
    if (!t.valid) throw std::runtime_error("rollout_backward: no rollout to differentiate");
    if (!gx) throw std::runtime_error("rollout_backward: null position cotangent");
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    ensure_tape();
    const int n = int(n_);
    const int every = t.every, S = t.substeps, G = (S + every - 1) / every;
    const int pp_grid = std::max(1, (n + 255) / 256);
    const int shape_grid = int(std::max<size_t>(1, std::min<size_t>((cmax_ + 255) / 256, 2048)));
    if (t.relax_part.n != size_t(pp_grid + shape_grid)) dalloc(t.relax_part, size_t(pp_grid + shape_grid));
    HIPCHECK(hipMemsetAsync(t.relax_part.p, 0, t.relax_part.n * sizeof(double), stream_));
    HIPCHECK(hipMemsetAsync(t.shx.p, 0, t.shx.n * sizeof(float), stream_));
    HIPCHECK(hipMemsetAsync(t.shv.p, 0, t.shv.n * sizeof(float), stream_));
    // keep the forward's final state (and its slot order)
    HIPCHECK(hipMemcpyAsync(t.finP.p, P_[s_].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(t.finV.p, V_[s_].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(t.finF.p, flags_cur(), n_ * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(t.finI.p, ids_cur(), n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
    int cur = 0;
    launch_fill_f4(t.lx[cur].p, n, stream_);
    launch_fill_f4(t.lv[cur].p, n, stream_);
    launch_gather_f3(t.lx[cur].p, gx, ids_cur(), n, stream_);   // final cotangents into the final slot order
    if (gv) launch_gather_f3(t.lv[cur].p, gv, ids_cur(), n, stream_);
    const float keep = relax_;
    relax_ = t.relax;
    const float inv_cell = 1.0f / cfg_.search_radius;
    try {
        for (int seg = G - 1; seg >= 0; seg--) {
            const int s0 = seg * every, s1 = std::min(S, s0 + every);
            // recompute the segment from its checkpoint, recording each substep
            HIPCHECK(hipMemcpyAsync(P_[s_].p, t.ckP[seg].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
            HIPCHECK(hipMemcpyAsync(V_[s_].p, t.ckV[seg].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
            HIPCHECK(hipMemcpyAsync(flags_cur(), t.ckF[seg].p, n_ * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
            HIPCHECK(hipMemcpyAsync(ids_cur(), t.ckI[seg].p, n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
            launch_frame_begin(cc_.p, mcnt_.p, int(mcnt_.n), stream_);
            for (int j = 0; j < s1 - s0; j++) {
                HIPCHECK(hipMemcpyAsync(t.P0[j].p, P_[s_].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
                HIPCHECK(hipMemcpyAsync(t.V0[j].p, V_[s_].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
                swap_tape(j);
                try {
                    substep(t.sdt, 1, true, j & 1, true);
                } catch (...) {
                    swap_tape(j);
                    throw;
                }
                swap_tape(j);
                HIPCHECK(hipMemcpyAsync(t.Fs[j].p, flags_cur(), n_ * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
                HIPCHECK(hipMemcpyAsync(t.counts.p + j, cc_.p + (j & 1), sizeof(int32_t), hipMemcpyDeviceToDevice,
                                        stream_));
            }
            // reverse the segment
            for (int j = s1 - s0 - 1; j >= 0; j--) {
                const int nxt = cur ^ 1;
                build_ranges(t.keys[j].p);
                AdjArgs a;
                a.n = n; a.keys = t.keys[j].p; a.vals = t.vals[j].p; a.range = range_.p; a.Fs = t.Fs[j].p;
                a.Xs = t.Xs[j].p; a.Vs = t.Vs[j].p; a.P0 = t.P0[j].p; a.V0 = t.V0[j].p; a.delta = t.delta[j].p;
                a.lx_in = t.lx[cur].p; a.lv_in = t.lv[cur].p; a.lx_out = t.lx[nxt].p; a.lv_out = t.lv[nxt].p;
                a.As = t.As.p; a.shx = t.shx.p; a.shv = t.shv.p;
                a.relax_part = t.relax_part.p;
                a.mu = cfg_.particle_mu; a.coh = cfg_.particle_cohesion; a.max_r = cfg_.particle_max_radius;
                a.dt = t.sdt; a.relax = t.relax; a.inv_cell = inv_cell; a.vmax = cfg_.particle_v_max;
                a.gx = cfg_.gravity[0]; a.gy = cfg_.gravity[1]; a.gz = cfg_.gravity[2];
                a.rmax = rmax_;
                a.planes = planes_ ? plane_bits_ : 0;
                a.rowtab = rowtab();
                launch_adj_prep(a, stream_);
                AdjShapeArgs sa;
                sa.cmax = int(cmax_); sa.cc = t.counts.p + j;
                sa.cpart = t.cpart[j].p; sa.cshape = t.cshape[j].p; sa.cbpos = t.cbpos[j].p; sa.cbvel = t.cbvel[j].p;
                sa.cnorm = t.cnorm[j].p; sa.flags = t.Fs[j].p; sa.Xs = t.Xs[j].p; sa.Vs = t.Vs[j].p;
                sa.As = t.As.p; sa.body = body_args();
                sa.pmu = cfg_.soft_contact_mu; sa.ka = cfg_.particle_adhesion; sa.dt = t.sdt; sa.relax = t.relax;
                sa.shx = t.shx.p; sa.shv = t.shv.p; sa.relax_part = t.relax_part.p + pp_grid;
                launch_adj_shape(sa, shape_grid, stream_);
                launch_adj_pp(a, num_scenes_ > 1, stream_);
                HIPCHECK(hipGetLastError());
                cur = nxt;
            }
        }
    } catch (...) {
        relax_ = keep;
        throw;
    }
    relax_ = keep;
    // start cotangents are in the slot order of the first checkpoint
    if (gx0) launch_scatter_f3(gx0, t.lx[cur].p, t.ckI[0].p, n, stream_);
    if (gv0) launch_scatter_f3(gv0, t.lv[cur].p, t.ckI[0].p, n, stream_);
    std::vector<double> parts(t.relax_part.n);
    HIPCHECK(hipMemcpyAsync(parts.data(), t.relax_part.p, parts.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipMemcpyAsync(P_[s_].p, t.finP.p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(V_[s_].p, t.finV.p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(flags_cur(), t.finF.p, n_ * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(ids_cur(), t.finI.p, n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    double sum = 0.0;
    for (double v : parts) sum += v;
    if (grelax) *grelax = sum;
}

}  // namespace xpbd
