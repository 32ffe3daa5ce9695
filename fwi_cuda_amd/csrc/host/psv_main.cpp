// psv_main.cpp -- fwi_psv: the reference's seis_fwi driver on the gfx950 library.
//
// Same inputs and outputs as the reference executable (h_main -> simulate_PSV,
// src/h_main.cpp:6-11, src/h_simulate_PSV.cpp:92-258): reads ./bin/metaint.bin,
// intarray.bin, metafloat.bin, mat.bin (and ./io/shot*_rtf_*.bin for FWI) from
// the working directory (or the directory given as the first argument), builds
// the CPML profiles, Holberg coefficients and source wavelets, then runs
// forward modelling (simulate_fwd_PSV) or FWI iterations (simulate_fwi_PSV)
// through the C ABI and writes the reference's output files.
//
// Differences from the reference, all explicit:
//   --niter N      FWI iterations (the reference's CPU loop has no working stop
//                  rule, src/cpu/n_simulate_PSV.cpp:576-581; its GPU twin stops at 10)
//   --grad-sign S  -1: the CPU path's gradient sign (default), +1: the GPU path's
//   --batch B      shots per launch (default: as many as memory allows)
//   --storage full|checkpoint|auto
//   --seis-fp64    seismograms, misfits and adjoint sources from an FP64 propagation
//                  (the reference's precision for them; the wavefields stay FP32)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "fwi_psv.h"
#include "psv_io.h"

namespace {

int die(const char *what, int status, const fwi_ctx *ctx) {
    std::fprintf(stderr, "fwi_psv: %s failed: %s (%s)\n", what, fwi_strerror(status), fwi_last_error(ctx));
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    std::string dir = ".";
    int niter = 10, grad_sign = -1, batch = 0, storage = FWI_STORE_AUTO, seis_fp64 = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "--niter") niter = std::atoi(next());
        else if (a == "--grad-sign") grad_sign = std::atoi(next());
        else if (a == "--batch") batch = std::atoi(next());
        else if (a == "--seis-fp64") seis_fp64 = 1;
        else if (a == "--storage") {
            std::string s = next();
            storage = s == "full" ? FWI_STORE_FULL : s == "checkpoint" ? FWI_STORE_CHECKPOINT : FWI_STORE_AUTO;
        } else if (a[0] != '-') dir = a;
    }
    std::string err;
    psvio::MetaInt m{};
    psvio::IntArrays ia;
    psvio::MetaFloat mf{};
    if (!psvio::read_metaint(dir, m, err) || !psvio::read_intarray(dir, m, ia, err) ||
        !psvio::read_metafloat(dir, mf, err)) {
        std::fprintf(stderr, "fwi_psv: %s\n", err.c_str());
        return 1;
    }
    const int nz = m.nz, nx = m.nx, nt = m.nt, nshot = m.nshot;
    const size_t n = (size_t)nz * nx;
    std::vector<double> lam, mu, rho;
    if (m.mat_grid) {
        if (!psvio::read_material(dir, nz, nx, lam, mu, rho, err)) {
            std::fprintf(stderr, "fwi_psv: %s\n", err.c_str());
            return 1;
        }
    } else {  // mat_stag_PSV (src/h_preproc_PSV.cpp:99-110)
        lam.assign(n, mf.scalar_lam);
        mu.assign(n, mf.scalar_mu);
        rho.assign(n, mf.scalar_rho);
    }
    // CPML profiles and Holberg coefficients (src/h_simulate_PSV.cpp:155-170)
    std::vector<double> pz(6 * (size_t)nz), px(6 * (size_t)nx);
    fwi_grid g{};
    g.nt = nt; g.nz = nz; g.nx = nx;
    g.dt = mf.dt; g.dz = mf.dz; g.dx = mf.dx;
    g.fdorder = m.fdorder;
    int st = fwi_holbergcoeff(m.fdorder, 1, g.hc);
    if (st) return die("holbergcoeff", st, nullptr);
    for (int i = 0; i < 4; ++i) g.isurf[i] = m.surf ? ia.isurf[i] : 0;
    g.pml_z = m.pml_z;
    g.pml_x = m.pml_x;
    double *z6[6], *x6[6];
    for (int k = 0; k < 6; ++k) {
        z6[k] = pz.data() + k * (size_t)nz;
        x6[k] = px.data() + k * (size_t)nx;
    }
    if (m.pml_z && (st = fwi_cpml(z6[0], z6[1], z6[2], z6[3], z6[4], z6[5], mf.npower, mf.damp_v_pml, mf.rcoef,
                                  mf.k_max_pml, mf.freq_pml, ia.npml[0], ia.npml[1], m.fpad, nz, mf.dt, mf.dz)))
        return die("cpml (z)", st, nullptr);
    if (m.pml_x && (st = fwi_cpml(x6[0], x6[1], x6[2], x6[3], x6[4], x6[5], mf.npower, mf.damp_v_pml, mf.rcoef,
                                  mf.k_max_pml, mf.freq_pml, ia.npml[2], ia.npml[3], m.fpad, nx, mf.dt, mf.dx)))
        return die("cpml (x)", st, nullptr);
    g.a_z = z6[0]; g.b_z = z6[1]; g.K_z = z6[2]; g.a_half_z = z6[3]; g.b_half_z = z6[4]; g.K_half_z = z6[5];
    g.a_x = x6[0]; g.b_x = x6[1]; g.K_x = x6[2]; g.a_half_x = x6[3]; g.b_half_x = x6[4]; g.K_half_x = x6[5];
    g.snap_z1 = m.snap_z1; g.snap_z2 = m.snap_z2; g.snap_x1 = m.snap_x1; g.snap_x2 = m.snap_x2;
    g.snap_dt = m.snap_dt; g.snap_dz = m.snap_dz; g.snap_dx = m.snap_dx;
    std::memcpy(g.taper, m.taper, sizeof g.taper);
    // stability check (checkfd_ssg_elastic), warning only as in the reference
    double cmax = 0, cmin = 0;
    const double dtstab = fwi_dtstab(nz, nx, mf.dz, ia.npml[1], rho.data(), lam.data(), mu.data(), g.hc, &cmax, &cmin);
    std::printf("nz = %d, nx = %d, dz = %g, dt = %g, DTSTAB = %g%s\n", nz, nx, mf.dz, mf.dt, dtstab,
                mf.dt > dtstab ? "  (unstable: choose a smaller dt)" : "");
    // wavelets: the reference fires an x force only (STF_Z amplitude 0, STF_X 1,
    // src/h_simulate_PSV.cpp:182-185); sources grouped by the shot they fire in
    std::vector<double> w(nt);
    fwi_wavelet(w.data(), nt, mf.dt, 1.0, mf.freq_pml, 0.0, 1);
    std::vector<int32_t> src_off(nshot + 1, 0), sz, sx, rec_off(nshot + 1), rz, rx;
    std::vector<double> stf_z, stf_x;
    for (int s = 0; s < nshot; ++s) {
        for (int i = 0; i < m.nsrc; ++i)
            if (ia.fire[i] == s) {
                sz.push_back(ia.z_src[i]);
                sx.push_back(ia.x_src[i]);
                stf_z.insert(stf_z.end(), nt, 0.0);
                stf_x.insert(stf_x.end(), w.begin(), w.end());
            }
        src_off[s + 1] = (int32_t)sz.size();
        rec_off[s] = s * m.nrec;
        rz.insert(rz.end(), ia.z_rec.begin(), ia.z_rec.end());
        rx.insert(rx.end(), ia.x_rec.begin(), ia.x_rec.end());
    }
    rec_off[nshot] = nshot * m.nrec;
    fwi_acq q{nshot, m.stf_type, src_off.data(), sz.data(), sx.data(), stf_z.data(), stf_x.data(),
              rec_off.data(), rz.data(), rx.data()};
    fwi_opts o;
    fwi_default_opts(&o);
    o.batch = batch;
    o.storage = (!m.fwinv && m.accu_save) ? FWI_STORE_FULL : storage;
    o.grad_sign = grad_sign;
    o.seis_fp64 = seis_fp64;
    fwi_ctx *ctx = nullptr;
    if ((st = fwi_create(&g, &q, &o, &ctx))) return die("create", st, nullptr);
    if ((st = fwi_set_model(ctx, lam.data(), mu.data(), rho.data()))) return die("set_model", st, ctx);
    const size_t rows = (size_t)m.nrec * nt;
    if (!m.fwinv) {
        // simulate_fwd_PSV (src/cpu/n_simulate_PSV.cpp:18-149)
        std::vector<double> uz(rows), ux(rows);
        int32_t snt = 1 + (nt - 1) / m.snap_dt;
        size_t cube = (size_t)snt * (1 + (m.snap_z2 - m.snap_z1) / m.snap_dz) * (1 + (m.snap_x2 - m.snap_x1) / m.snap_dx);
        std::vector<double> acc(m.accu_save ? 5 * cube : 0);
        for (int s = 0; s < nshot; ++s) {
            std::printf("FORWARD KERNEL: SHOT %d of %d.\n", s, nshot);
            if (m.accu_save) {
                double *f[5];
                for (int k = 0; k < 5; ++k) f[k] = acc.data() + k * cube;
                if ((st = fwi_forward_snapshots(ctx, s, uz.data(), ux.data(), f[0], f[1], f[2], f[3], f[4])))
                    return die("forward", st, ctx);
                psvio::write_accu(dir, s, cube, f);
            } else if ((st = fwi_forward(ctx, s, 1, uz.data(), ux.data()))) {
                return die("forward", st, ctx);
            }
            if (m.seismo_save) psvio::write_seismo(dir, s, m.nrec, nt, uz.data(), ux.data());
        }
    } else {
        // simulate_fwi_PSV (src/cpu/n_simulate_PSV.cpp:153-597)
        if (!m.rtf_meas_true) {
            std::fprintf(stderr, "fwi_psv: no field measurement exists, unable to run FWI\n");
            return 1;
        }
        std::vector<double> oz(rows * nshot), ox(rows * nshot);
        for (int s = 0; s < nshot; ++s)
            if (!psvio::read_seismo(dir, s, m.nrec, nt, oz.data() + s * rows, ox.data() + s * rows, err)) {
                std::fprintf(stderr, "fwi_psv: %s\n", err.c_str());
                return 1;
            }
        if ((st = fwi_set_observed(ctx, oz.data(), ox.data()))) return die("set_observed", st, ctx);
        std::vector<double> L2s, mis(nshot);
        double est = 0.01;   // first step-length estimate (src/cpu/n_simulate_PSV.cpp:243)
        for (int it = 0; it < niter; ++it) {
            // one simulate_fwi_PSV iteration (src/cpu/n_simulate_PSV.cpp:249-582) through the ABI:
            // copy_mat, mat_av2 of the current model, the batched shot gradients, taper2,
            // step_length_PSV and update_mat2
            std::printf("FWI: Iteration %d\n", it);
            if ((st = fwi_save_model(ctx))) return die("copy_mat", st, ctx);
            if ((st = fwi_get_model(ctx, lam.data(), mu.data(), rho.data())) ||
                (st = fwi_set_model(ctx, lam.data(), mu.data(), rho.data())))
                return die("mat_av2", st, ctx);
            if ((st = fwi_gradient(ctx, 0, nshot, mis.data()))) return die("gradient", st, ctx);
            double L2 = 0;
            for (int s = 0; s < nshot; ++s) {
                L2 += mis[s];
                std::printf("L2 NORM: %d, %.9g\n", s, L2);
            }
            if ((st = fwi_taper_gradient(ctx))) return die("taper", st, ctx);
            double step = 0;
            int32_t ntrials = 0;
            st = fwi_step_length(ctx, est, L2, &step, &ntrials);
            if (st == FWI_ESTEP) {   // the reference prints this and exit(0)s (n_step_length_PSV.cpp:173-178)
                std::printf("Steplength estimation failed! (%d trials)\n", ntrials);
                fwi_destroy(ctx);
                return 2;
            }
            if (st) return die("step length", st, ctx);
            if ((st = fwi_update_model(ctx, step, 0.5, 1.25))) return die("update_mat2", st, ctx);
            est = step;
            L2s.push_back(L2);
            std::printf("FWI: Iteration %d  L2 = %.9g  step = %.6g  (%d trials)\n", it, L2, step, ntrials);
            if (m.mat_save_interval > 0 && it % m.mat_save_interval == 0) {
                fwi_get_model(ctx, lam.data(), mu.data(), rho.data());
                psvio::write_mat(dir, it, n, lam.data(), mu.data(), rho.data());
            }
        }
        if (m.mat_save_interval < 1) {
            fwi_get_model(ctx, lam.data(), mu.data(), rho.data());
            psvio::write_mat(dir, niter, n, lam.data(), mu.data(), rho.data());
        }
        std::printf("L2 norms: ");
        for (double v : L2s) std::printf("%.9g, ", v);
        std::printf("\n");
    }
    fwi_destroy(ctx);
    std::printf("Simulation Complete\n");
    return 0;
}
