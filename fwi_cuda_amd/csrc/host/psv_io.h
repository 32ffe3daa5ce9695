// psv_io.h -- the reference's binary file layout (all little-endian, no headers).
//
//   ./bin/metaint.bin    37 int32     src/h_io_PSV.cpp:19-107
//   ./bin/intarray.bin   int32 lists  src/h_io_PSV.cpp:109-164
//   ./bin/metafloat.bin  11 float64   src/h_io_PSV.cpp:167-199
//   ./bin/mat.bin        float64 lam, mu, rho [nz][nx]        :201-242
//   ./io/shot{i}_rtf_u{z,x}.bin   observed [nrec][nt] float64   :320-345
//   ./bin/shot{i}_rtf_u{z,x}.bin  modelled [nrec][nt]           :292-317
//   ./bin/shot{i}_{vz,vx,szz,szx,sxx}.bin  [snap_nt][snap_nz][snap_nx]  :246-288
//   ./bin/iter{k}_mat.bin         model after iteration k       :377-412
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace psvio {

struct MetaInt {
    int gpu, surf, pml_z, pml_x, accu_save, seismo_save, fwinv, rtf_meas_true;
    int nt, nz, nx, snap_t1, snap_t2, snap_z1, snap_z2, snap_x1, snap_x2, snap_dt, snap_dz, snap_dx;
    int taper[8];
    int nsrc, nrec, nshot, stf_type, rtf_type, fdorder, fpad, mat_save_interval, mat_grid;
};

struct IntArrays {
    int npml[4], isurf[4];
    std::vector<int32_t> z_src, x_src, fire, z_rec, x_rec;
};

struct MetaFloat {
    double dt, dz, dx, npower, damp_v_pml, rcoef, k_max_pml, freq_pml, scalar_lam, scalar_mu, scalar_rho;
};

bool read_metaint(const std::string &dir, MetaInt &m, std::string &err);
bool read_intarray(const std::string &dir, const MetaInt &m, IntArrays &a, std::string &err);
bool read_metafloat(const std::string &dir, MetaFloat &f, std::string &err);
bool read_material(const std::string &dir, int nz, int nx, std::vector<double> &lam, std::vector<double> &mu,
                   std::vector<double> &rho, std::string &err);
bool read_seismo(const std::string &dir, int ishot, int nrec, int nt, double *uz, double *ux, std::string &err);
bool write_seismo(const std::string &dir, int ishot, int nrec, int nt, const double *uz, const double *ux);
bool write_accu(const std::string &dir, int ishot, size_t n, const double *const fields[5]);
bool write_mat(const std::string &dir, int iter, size_t n, const double *lam, const double *mu, const double *rho);

}  // namespace psvio
