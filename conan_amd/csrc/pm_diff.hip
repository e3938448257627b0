// pm_diff.hip -- the differentiable PM driver of include/conanpm.h
// (pm_diff_*): white noise -> scale-free IC -> Zel'dovich start -> leapfrog
// in the FP32 form -> final CIC density, and its reverse-mode adjoint.
//
// Forward: the same device operations as pm_ic_scale_free + pm_init_zeldovich
// + pm_steps + pm_deposit (so the density is bit-identical to that sequence),
// storing the particle state (positions, momenta, ids) and its checksum
// before every ckpt_every-th kick, and with PM_DIFF_STORE_POTENTIALS the
// potential of every kick.
//
// Backward, from lY = m dW/dY . dL/drho at the final positions:
//   for i = n-1 .. 0:  sort, phi_i (stored, or deposit + Poisson:
//     bit-identical to the forward), deposit of the momentum adjoint, -div,
//     Poisson (chi), fused adjoint gather that undoes kick i and (i > 0)
//     the drift before it (exact) with the drift's adjoint; at a checkpoint
//     the same pass takes the checksum of the state between the two undos
//     and compares it with the stored one; on a mismatch, or with
//     PM_DIFF_VERIFY_CHECKPOINTS, the drift is redone on the positions, the
//     state compared value by value and the pass continues from the stored
//     state.
//   then the Zel'dovich start and the IC field in FP64 on the mass grid:
//     X = q - a0 u, P = -a_mom u, u = D phi_IC / res  (src/run/nbody.hh:40-50)
//     phi_IC = Pi delta, Pi = IFFT[-1/k^2 FFT] / n       (src/ic/ic.cc:100-117)
//     delta = sigma A w / sqrt(var), var = mean((B w)^2) (src/ic/ic.cc:59-82)
//   with A = IFFT[f_A FFT]/n, B = IFFT[f_B FFT]/n, f_B = sqrt(P) S and f_A =
//   f_B (smooth) or sqrt(P); A, B and Pi are symmetric (real even filters):
//     l_delta = Pi(-sum_k D_k l_u),  l_u = -a0 l_X - a_mom l_P
//     l_w = IFFT[ (sigma/sqrt(var)) f_A FFT(l_delta)
//                 - (<l_delta, delta> / (n var)) f_B^2 FFT(w) ] / n
#include "pm_adjoint.hpp"
#include "pm_impl.hpp"

namespace cpm {

struct DiffPM {
    pm_diff_params p;
    int R, lgNm, Nm;
    long Np;
    Impl<PosFX> sim;
    hipStream_t st;
    FFTPlan<double> icfft;
    DevBuf<double> noise, delta, icm, icred, icscal;
    DevBuf<cx<double>> ispec, inyq, ispec2, inyq2;
    double var = 0;
    std::vector<double> a_seq;  // a_i, accumulated like LeapFrog::run (a += da)
    int nck;
    std::vector<std::unique_ptr<DevBuf<uint32_t>>> ckX, ckid;
    std::vector<std::unique_ptr<DevBuf<double>>> ckP;
    DevBuf<unsigned long long> ckhash, hbuf;        // checkpoint checksums, scratch
    std::vector<std::unique_ptr<DevBuf<float>>> phis;  // PM_DIFF_STORE_POTENTIALS: phi of every kick
    bool store_phi, verify;
    DevBuf<float> lY, lP, lY2, lP2;  // adjoint state (FP32), in the storage order and the next
    DevBuf<uint32_t> tmpY, bmax, lmax;
    DevBuf<float> svec, psi, chi, grho;
    DevBuf<double> dpart, dscal;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool have_forward = false;
    pm_diff_stats stats{};

    static pm_config sim_config(const pm_diff_params& q) {
        pm_config c{};
        c.dim = q.dim;
        c.Nm = 1 << q.mbits;
        c.Nf = q.f << q.mbits;
        c.L = q.L;
        c.H0 = q.H0;
        c.Omega_m = q.Omega_m;
        c.Omega_L = q.Omega_L;
        c.precision = PM_FP32;
        // the differentiable map gathers on folded coordinates: its force is
        // a smooth function of the position on the torus (the adjoint kernels
        // stage the tile of the folded origin)
        c.flags = PM_FLAG_WRAP_GATHER;
        return c;
    }

    explicit DiffPM(const pm_diff_params& q)
        : p(q), R(q.dim), lgNm(q.mbits), Nm(1 << q.mbits), Np(1L << (q.dim * q.mbits)), sim(sim_config(q)),
          st(sim.st), icfft(q.dim, 1 << q.mbits), nck((q.nsteps + q.ckpt_every - 1) / q.ckpt_every),
          store_phi(q.flags & PM_DIFF_STORE_POTENTIALS), verify(q.flags & PM_DIFF_VERIFY_CHECKPOINTS) {
        size_t f0 = 0, t0 = 0, f1 = 0;
        CPM_HIP(hipMemGetInfo(&f0, &t0));
        noise.alloc(Np);
        delta.alloc(Np);
        icm.alloc(Np);
        icred.alloc(4096);
        icscal.alloc(4);
        ispec.alloc(icfft.spec_elems);
        inyq.alloc(icfft.nyq_elems);
        ispec2.alloc(icfft.spec_elems);
        inyq2.alloc(icfft.nyq_elems);
        for (int c = 0; c < nck; ++c) {
            ckX.emplace_back(new DevBuf<uint32_t>((size_t)R * Np));
            ckP.emplace_back(new DevBuf<double>((size_t)R * Np));
            ckid.emplace_back(new DevBuf<uint32_t>(Np));
        }
        ckhash.alloc(nck);
        hbuf.alloc(1);
        if (store_phi)
            for (int i = 0; i < q.nsteps; ++i) phis.emplace_back(new DevBuf<float>(sim.M));
        lY.alloc((size_t)R * Np);
        lP.alloc((size_t)R * Np);
        lY2.alloc((size_t)R * Np);
        lP2.alloc((size_t)R * Np);
        tmpY.alloc((size_t)R * Np);
        bmax.alloc(std::max((Np + 255) / 256, brick_count(sim.pa)));
        lmax.alloc(1);
        svec.alloc((size_t)R * sim.M);
        psi.alloc(sim.M);
        chi.alloc(sim.M);
        grho.alloc(sim.M);
        dpart.alloc(3 * 1024);
        dscal.alloc(4);
        CPM_HIP(hipEventCreate(&ev0));
        CPM_HIP(hipEventCreate(&ev1));
        CPM_HIP(hipMemGetInfo(&f1, &t0));
        // the simulation state was allocated before f0; count it separately
        const double sim_bytes = (double)Np * (2 * R * 4 + 2 * R * 8 + 2 * 4 + 4 * 4) +
                                 (double)sim.M * 4 + (double)(sim.fft.spec_elems + sim.fft.nyq_elems) * 8;
        stats.device_bytes = (double)(f0 - f1) + sim_bytes;
        stats.ncheckpoints = nck;
        a_seq.resize(q.nsteps + 1);
        double a = q.a0;
        for (int i = 0; i <= q.nsteps; ++i) {
            a_seq[i] = a;
            a += q.da;
        }
    }
    ~DiffPM() {
        (void)hipStreamSynchronize(st);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }

    FilterArgs ic_filter(int kind) const {
        FilterArgs f{};
        f.kind = kind;
        f.dim = R;
        f.lgN = lgNm;
        f.kscale = 2 * M_PI / (double)Nm;  // kspace<R>(N, N), src/ic/ic.cc:64
        f.ns = p.ns;
        f.t = p.scale * Nm / p.L;  // Fourier::scale<R>(sigma * N/L), ic.cc:63
        return f;
    }
    void ic_apply(cx<double>* s, cx<double>* n, const FilterArgs& f) {
        apply_filter(s, n, (long)icfft.spec_elems, (long)icfft.nyq_elems, f, st);
    }
    // in-place _compute_potential on icm (src/ic/ic.cc:100-117)
    void ic_potential() {
        icfft.forward(icm.p, ispec.p, inyq.p, 0.0, nullptr, st);
        FilterArgs f{};
        f.kind = FILT_POT;
        f.dim = R;
        f.lgN = lgNm;
        f.kscale = 2 * M_PI / p.L;
        ic_apply(ispec.p, inyq.p, f);
        icfft.backward(ispec.p, inyq.p, icm.p, (double)Np, st);
    }

    // Poisson of an arbitrary mesh (no mean subtraction: the DC mode is zeroed).
    void poisson_into(const float* in, float* out, double a) {
        const double s = sim.fft.M / sim.cos.G * a;  // the scale of Impl::do_poisson
        sim.fft.poisson(in, out, sim.spec.p, sim.nyq.p, 0.f, nullptr, 2 * M_PI / p.L, s, st);
    }

    void forward(const double* w, bool dev, double* rho, bool dev_out) {
        CPM_HIP(hipEventRecord(ev0, st));
        CPM_HIP(hipMemcpyAsync(noise.p, w, sizeof(double) * Np, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               st));
        // scale-free field: exactly the sequence of pm_ic_scale_free
        icfft.forward(noise.p, ispec.p, inyq.p, 0.0, nullptr, st);
        CPM_HIP(hipMemcpyAsync(ispec2.p, ispec.p, sizeof(cx<double>) * icfft.spec_elems, hipMemcpyDeviceToDevice, st));
        CPM_HIP(hipMemcpyAsync(inyq2.p, inyq.p, sizeof(cx<double>) * icfft.nyq_elems, hipMemcpyDeviceToDevice, st));
        ic_apply(ispec.p, inyq.p, ic_filter(FILT_PS));
        icfft.backward(ispec.p, inyq.p, icm.p, (double)Np, st);
        reduce_sum(icm.p, Np, 1, 1.0 / (double)Np, icred.p, 4096, icscal.p, st);
        CPM_HIP(hipMemcpyAsync(&var, icscal.p, sizeof(double), hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        if (!(var > 0)) throw Error(PM_EINVAL, "the smoothed noise field has zero variance");
        ic_apply(ispec2.p, inyq2.p, ic_filter(p.smooth ? FILT_PS : FILT_P));
        icfft.backward(ispec2.p, inyq2.p, delta.p, Np * sqrt(var) / p.sigma, st);  // ic.cc:80
        CPM_HIP(hipMemcpyAsync(icm.p, delta.p, sizeof(double) * Np, hipMemcpyDeviceToDevice, st));
        ic_potential();
        // Zel'dovich start and the leapfrog (pm_init_zeldovich + pm_steps)
        ParticleOps<PosFX>::zeldovich(R, icm.p, lgNm, p.L, p.a0, p.a0 + p.da / 2, sim.X.p, sim.P.p, st);
        sim.fx_quantize_momenta();
        iota_u32(sim.id.p, Np, st);
        sim.invalidate_order();  // storage order is the mass-grid order now
        sim.have_state = true;
        ParticleOps<PosFX>::drift(sim.X.p, sim.P.p, (long)R * Np, sim.drift_factor(a_seq[0], p.da), st);
        for (int i = 0; i < p.nsteps; ++i) {
            if (i % p.ckpt_every == 0) {  // copy and checksum in one pass
                const int c = i / p.ckpt_every;
                state_hash(R, sim.X.p, sim.P.p, sim.id.p, Np, ckX[c]->p, ckP[c]->p, ckid[c]->p, ckhash.p + c, st);
            }
            const bool more = i + 1 < p.nsteps;
            sim.phi_out = store_phi ? phis[i]->p : nullptr;
            sim.kick_impl(a_seq[i] + p.da / 2, p.da, more, a_seq[i + 1], p.da);
        }
        sim.phi_out = nullptr;
        // final density: the deposit of pm_deposit, with the fixed-point
        // accumulators converted to FP64 instead of rounded to FP32
        sort_particles();
        DevBuf<double> t;
        double* out = rho;
        if (!dev_out) {
            t.alloc(sim.M);
            out = t.p;
        }
        deposit_fx_brick<double>(R, sim.bstart, sim.spids, sim.X.p, Np, sim.lgNf, sim.pa, (float)sim.mass, out,
                                 sim.row_partial.p, st);
        if (!dev_out) {
            CPM_HIP(hipMemcpyAsync(rho, t.p, sizeof(double) * sim.M, hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
        }
        CPM_HIP(hipEventRecord(ev1, st));
        CPM_HIP(hipStreamSynchronize(st));
        float ms = 0;
        CPM_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        stats.forward_ms = ms;
        have_forward = true;
    }

    // keys of the current positions -> sort -> brick offsets (the first half
    // of Impl::do_deposit)
    void sort_particles() {
        sim.fx_sort_current();
    }

    void swap_all() {
        std::swap(sim.X.p, sim.X2.p);
        std::swap(sim.P.p, sim.P2.p);
        std::swap(sim.id.p, sim.id2.p);
        std::swap(lY.p, lY2.p);
        std::swap(lP.p, lP2.p);
    }

    // Drift factor of the adjoint, lP += cd lY, for the drift at a (lY in
    // 1/length): da / (a^2 adot), src/run/nbody.hh:124,129.
    double drift_adj(double a) const { return p.da / (a * a * sim.cos.adot(a)); }

    // Adjoint gather (undoes the kick).  fuse: also undo the drift before it
    // (cd_fx, cd; both 0 when there is none) and produce the next sort's keys
    // and max |lP| (into lmax).  hash != nullptr: the checksum of the state
    // between the two undos, into *hash.
    void adj_gather(const float* phi, const float* chiv, AdjCoef c, bool fuse, double cd_fx, double cd,
                    unsigned long long* hash = nullptr) {
        AdjDrift dr{0.0, 0.0, 0, nullptr, nullptr, nullptr, nullptr, hash};
        if (fuse) {
            zero_u32(sim.cnt3.p, 3 * sim.nsort_groups, st);
            sim.counts_clear = false;  // the adjoint gather counts into them
            dr = AdjDrift{cd_fx, cd, (cd_fx != 0.0 || cd != 0.0) ? 1 : 0, sim.knew.p, sim.cnt3.p,
                          sim.cnt3.p + sim.nsort_groups, bmax.p, hash};
        }
        adj_gather_fx_brick(R, sim.bstart, sim.spids, sim.X.p, sim.P.p, lY.p, lP.p, sim.id.p, sim.X2.p, sim.P2.p,
                            lY2.p, lP2.p, sim.id2.p, Np, phi, chiv, sim.lgNf, sim.pa, c, dr, lmax.p, st);
        swap_all();
        sim.resort_ok = true;  // the particles now sit in the order of the last sort
        sim.keys_ready = fuse;
    }

    // Whether the checksum in hbuf equals checkpoint c's: the reversed state
    // is then the checkpoint's up to the storage order, which nothing
    // downstream depends on.
    bool hash_matches(int c) {
        unsigned long long hv[2] = {0, 1};
        CPM_HIP(hipMemcpyAsync(&hv[0], hbuf.p, sizeof(hv[0]), hipMemcpyDeviceToHost, st));
        CPM_HIP(hipMemcpyAsync(&hv[1], ckhash.p + c, sizeof(hv[1]), hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        return hv[0] == hv[1];
    }

    // Compare the reversed state with checkpoint c value by value and continue
    // from the checkpoint: the adjoints are moved into its particle order via
    // the ids.
    void restore(int c) {
        const DevBuf<uint32_t>& cX = *ckX[c];
        const DevBuf<double>& cP = *ckP[c];
        const DevBuf<uint32_t>& cid = *ckid[c];
        // reversed state in the checkpoint's order: X2 / P2
        DevBuf<double> tP((size_t)R * Np);
        scatter_by_id<uint32_t>(sim.X.p, sim.id.p, Np, R, tmpY.p, st);
        gather_by_id<uint32_t>(tmpY.p, cid.p, Np, R, sim.X2.p, st);
        scatter_by_id<double>(sim.P.p, sim.id.p, Np, R, tP.p, st);
        gather_by_id<double>(tP.p, cid.p, Np, R, sim.P2.p, st);
        const int nb = 1024;
        state_diff(sim.X2.p, cX.p, sim.P2.p, cP.p, (long)R * Np, dpart.p, nb, st);
        std::vector<double> h(3 * nb);
        CPM_HIP(hipMemcpyAsync(h.data(), dpart.p, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st));
        // adjoints into the checkpoint's order
        scatter_by_id<float>(lY.p, sim.id.p, Np, R, lY2.p, st);
        gather_by_id<float>(lY2.p, cid.p, Np, R, lY.p, st);
        scatter_by_id<float>(lP.p, sim.id.p, Np, R, lP2.p, st);
        gather_by_id<float>(lP2.p, cid.p, Np, R, lP.p, st);
        CPM_HIP(hipMemcpyAsync(sim.X.p, cX.p, sizeof(uint32_t) * R * Np, hipMemcpyDeviceToDevice, st));
        CPM_HIP(hipMemcpyAsync(sim.P.p, cP.p, sizeof(double) * R * Np, hipMemcpyDeviceToDevice, st));
        CPM_HIP(hipMemcpyAsync(sim.id.p, cid.p, sizeof(uint32_t) * Np, hipMemcpyDeviceToDevice, st));
        sim.invalidate_order();  // the checkpoint's order is not the last sort's
        CPM_HIP(hipStreamSynchronize(st));
        double cnt = 0, mdp = 0, mp = 0;
        for (int b = 0; b < nb; ++b) {
            cnt += h[3 * b];
            mdp = std::max(mdp, h[3 * b + 1]);
            mp = std::max(mp, h[3 * b + 2]);
        }
        stats.recon_pos += (long long)cnt;
        stats.recon_mom = std::max(stats.recon_mom, mp > 0 ? mdp / mp : mdp);
    }

    void backward(const double* g, bool dev, double* gw, bool dev_out) {
        if (!have_forward) throw Error(PM_ESTATE, "pm_diff_backward needs a pm_diff_forward first");
        have_forward = false;  // the backward pass consumes the final state
        stats.recon_pos = 0;
        stats.recon_mom = 0;
        CPM_HIP(hipEventRecord(ev0, st));
        {
            DevBuf<double> t;
            const double* src = g;
            if (!dev) {
                t.alloc(sim.M);
                CPM_HIP(hipMemcpyAsync(t.p, g, sizeof(double) * sim.M, hipMemcpyHostToDevice, st));
                src = t.p;
            }
            f64_to_f32(src, sim.M, grho.p, st);
            CPM_HIP(hipStreamSynchronize(st));
        }
        CPM_HIP(hipMemsetAsync(lY.p, 0, sizeof(float) * R * Np, st));
        CPM_HIP(hipMemsetAsync(lP.p, 0, sizeof(float) * R * Np, st));
        const double m = sim.mass;
        // final density: rho = m sum W(Y, c)  ->  lY = m dW/dY . grad_rho;
        // the last kick has no drift after it: only the keys and max |lP|.
        // The forward's final sort still describes the particles.
        adj_gather(nullptr, grho.p, AdjCoef{0.0, 0.0, m, sim.pq, 1.0 / sim.pq}, true, 0.0, 0.0);
        bool drifted = true;      // the drift after kick i is undone (keys, lmax ready)
        bool lam_drifted = false;  // its adjoint is already in lP (a checkpoint restored after a fused undo)
        for (int i = p.nsteps - 1; i >= 0; --i) {
            const double ak = a_seq[i] + p.da / 2;
            const double ck = p.da / sim.cos.adot(ak);  // src/run/nbody.hh:112-116
            if (!drifted) {
                // the drift after kick i (at a_{i+1}, fused into kick i forward)
                const Impl<PosFX>::FxKeyArgs ka = sim.fx_key_args();
                rev_drift_keys(R, sim.X.p, sim.P.p, lY.p, lP.p, Np, sim.drift_factor(a_seq[i + 1], p.da),
                               lam_drifted ? 0.0 : drift_adj(a_seq[i + 1]), sim.lgNf, sim.pa, ka.prevkeys, ka.keys,
                               ka.pids, ka.in_count, ka.out_count, bmax.p, lmax.p, st);
                lam_drifted = false;
            }
            sim.fx_sort_current();
            const float* phi = nullptr;
            if (store_phi) {
                phi = phis[i]->p;
                deposit_fx_vec(R, sim.bstart, sim.spids, sim.X.p, lP.p, Np, sim.lgNf, sim.pa, lmax.p, svec.p, st);
            } else {
                // the deposit of lP and the mass deposit of the same positions
                // in one launch (the mesh of deposit_fx_brick; the Poisson
                // solve needs no mesh sum)
                deposit_fx_vm(R, sim.bstart, sim.spids, sim.X.p, lP.p, Np, sim.lgNf, sim.pa, lmax.p, svec.p,
                              (float)m, sim.mesh.p, nullptr, st);
                sim.npartials = 0;
                sim.do_poisson(ak);  // phi_i, as in the forward
                phi = sim.mesh.p;
            }
            neg_div_f32(R, svec.p, sim.lgNf, sim.pa.res, psi.p, st);
            poisson_into(psi.p, chi.p, ak);
            const AdjCoef coef{ck, -ck, -ck * m, sim.pq, 1.0 / sim.pq};
            if (i % p.ckpt_every == 0) {
                // the checkpoint holds the state before kick i: its checksum is
                // taken in the pass that undoes kick i and the drift before it
                const double cd_fx = i > 0 ? sim.drift_factor(a_seq[i], p.da) : 0.0;
                adj_gather(phi, chi.p, coef, true, cd_fx, i > 0 ? drift_adj(a_seq[i]) : 0.0,
                           verify ? nullptr : hbuf.p);
                drifted = true;
                if (verify || !hash_matches(i / p.ckpt_every)) {
                    // redo the drift on the positions (Y + round(P cd_fx), exact:
                    // rounding to nearest is odd), compare value by value and
                    // continue from the checkpoint; lP already holds the drift's
                    // adjoint
                    const Impl<PosFX>::FxKeyArgs ka = sim.fx_key_args();
                    rev_drift_keys(R, sim.X.p, sim.P.p, lY.p, lP.p, Np, -cd_fx, 0.0, sim.lgNf, sim.pa, ka.prevkeys,
                                   ka.keys, ka.pids, ka.in_count, ka.out_count, bmax.p, lmax.p, st);
                    restore(i / p.ckpt_every);
                    drifted = false;
                    lam_drifted = true;
                }
            } else {
                adj_gather(phi, chi.p, coef, true, sim.drift_factor(a_seq[i], p.da), drift_adj(a_seq[i]));
                drifted = true;
            }
        }
        sim.have_deposit = false;
        // Zel'dovich start: Y0 = X0 + cd0 P0, X0 = q - a0 u, P0 = -a_mom u
        const double a0 = a_seq[0], amom = p.a0 + p.da / 2;
        const double cd0 = p.da / (a0 * a0 * sim.cos.adot(a0));
        // (the state's momenta are consumed: P2 and P take l_u)
        axpby_f32((long)R * Np, -a0 - amom * cd0, lY.p, -amom, lP.p, sim.P2.p, st);  // l_u, storage order
        scatter_by_id<double>(sim.P2.p, sim.id.p, Np, R, sim.P.p, st);              // l_u, mass-grid order
        neg_div_f64(R, sim.P.p, lgNm, p.L / Nm, icm.p, st);                          // l_phi
        ic_potential();                                                        // l_delta (Pi symmetric)
        dot(icm.p, delta.p, Np, icred.p, 4096, icscal.p, st);
        double ld = 0;
        CPM_HIP(hipMemcpyAsync(&ld, icscal.p, sizeof(double), hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        const double alpha = p.sigma / sqrt(var);
        const double beta = -ld / ((double)Np * var);  // 2 l_var / n with l_var = -<l_delta, delta> / (2 var)
        icfft.forward(icm.p, ispec.p, inyq.p, 0.0, nullptr, st);
        ic_apply(ispec.p, inyq.p, ic_filter(p.smooth ? FILT_PS : FILT_P));
        icfft.forward(noise.p, ispec2.p, inyq2.p, 0.0, nullptr, st);
        ic_apply(ispec2.p, inyq2.p, ic_filter(FILT_PS));
        ic_apply(ispec2.p, inyq2.p, ic_filter(FILT_PS));
        axpby(2L * (long)icfft.spec_elems, alpha, reinterpret_cast<double*>(ispec.p), beta,
              reinterpret_cast<double*>(ispec2.p), reinterpret_cast<double*>(ispec.p), st);
        axpby(2L * (long)icfft.nyq_elems, alpha, reinterpret_cast<double*>(inyq.p), beta,
              reinterpret_cast<double*>(inyq2.p), reinterpret_cast<double*>(inyq.p), st);
        if (dev_out) {
            icfft.backward(ispec.p, inyq.p, gw, (double)Np, st);
        } else {
            icfft.backward(ispec.p, inyq.p, icm.p, (double)Np, st);
            CPM_HIP(hipMemcpyAsync(gw, icm.p, sizeof(double) * Np, hipMemcpyDeviceToHost, st));
        }
        CPM_HIP(hipEventRecord(ev1, st));
        CPM_HIP(hipStreamSynchronize(st));
        float ms = 0;
        CPM_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        stats.backward_ms = ms;
    }
};

}  // namespace cpm

struct pm_diff {
    std::unique_ptr<cpm::DiffPM> impl;
    std::string err;
};

static thread_local std::string g_diff_err;

#define PMD_TRY try {
#define PMD_CATCH(D)                                                    \
    }                                                                   \
    catch (const cpm::Error& e) {                                       \
        if (D) (D)->err = e.what(); else g_diff_err = e.what();          \
        return e.code;                                                  \
    }                                                                   \
    catch (const std::exception& e) {                                   \
        if (D) (D)->err = e.what(); else g_diff_err = e.what();          \
        return PM_EDEVICE;                                              \
    }                                                                   \
    return PM_OK;

extern "C" {

int pm_diff_create(const pm_diff_params* p, pm_diff** out) {
    pm_diff* d = nullptr;
    PMD_TRY
    if (!p || !out) throw cpm::Error(PM_EINVAL, "null argument");
    *out = nullptr;
    if (p->dim != 2 && p->dim != 3) throw cpm::Error(PM_EUNSUP, "dim must be 2 or 3");
    if (p->mbits < 2 || p->mbits > 10) throw cpm::Error(PM_EUNSUP, "mbits must lie in [2, 10]");
    if (p->f < 1 || (p->f & (p->f - 1)) || (p->f << p->mbits) > 1024)
        throw cpm::Error(PM_EUNSUP, "f must be a power of two with f * 2^mbits <= 1024");
    if (p->nsteps < 1 || p->ckpt_every < 1) throw cpm::Error(PM_EINVAL, "nsteps and ckpt_every must be >= 1");
    if (std::pow((double)p->f, p->dim) >= 8388608.0)
        throw cpm::Error(PM_EUNSUP, "particle mass f^dim must be below 2^23 (FP32 deposit quantum)");
    if (p->precision != PM_FP32) throw cpm::Error(PM_EUNSUP, "the adjoint is implemented for the FP32 form");
    if (p->flags & ~(PM_DIFF_STORE_POTENTIALS | PM_DIFF_VERIFY_CHECKPOINTS))
        throw cpm::Error(PM_EINVAL, "unknown pm_diff flags");
    if (!(p->L > 0) || !(p->a0 > 0) || !(p->da > 0) || !(p->sigma > 0) || !(p->H0 > 0))
        throw cpm::Error(PM_EINVAL, "L, a0, da, sigma and H0 must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw cpm::Error(PM_EDEVICE, "no HIP device available");
    std::unique_ptr<pm_diff> h(new pm_diff);
    h->impl.reset(new cpm::DiffPM(*p));
    *out = h.release();
    PMD_CATCH(d)
}

void pm_diff_destroy(pm_diff* d) { delete d; }

const char* pm_diff_last_error(const pm_diff* d) { return d ? d->err.c_str() : g_diff_err.c_str(); }

int pm_diff_forward(pm_diff* d, const double* noise, double* rho, int device) {
    if (!d) {
        g_diff_err = "null handle";
        return PM_EINVAL;
    }
    PMD_TRY
    if (!noise || !rho) throw cpm::Error(PM_EINVAL, "null argument");
    if (!device) cpm::check_finite(noise, d->impl->Np, "noise");
    d->impl->forward(noise, device != 0, rho, device != 0);
    PMD_CATCH(d)
}

int pm_diff_backward(pm_diff* d, const double* grad_rho, double* grad_noise, int device) {
    if (!d) {
        g_diff_err = "null handle";
        return PM_EINVAL;
    }
    PMD_TRY
    if (!grad_rho || !grad_noise) throw cpm::Error(PM_EINVAL, "null argument");
    if (!device) cpm::check_finite(grad_rho, d->impl->sim.M, "grad_rho");
    d->impl->backward(grad_rho, device != 0, grad_noise, device != 0);
    PMD_CATCH(d)
}

int pm_diff_get_stats(const pm_diff* d, pm_diff_stats* s) {
    if (!d || !s) {
        g_diff_err = "null argument";
        return PM_EINVAL;
    }
    *s = d->impl->stats;
    return PM_OK;
}

}  // extern "C"
