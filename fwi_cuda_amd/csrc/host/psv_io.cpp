// psv_io.cpp -- readers and writers for the reference's input and output files.
#include "psv_io.h"

#include <fstream>

namespace psvio {

namespace {

template <class T>
bool read_all(const std::string &path, T *dst, size_t n, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    f.read(reinterpret_cast<char *>(dst), (std::streamsize)(n * sizeof(T)));
    if ((size_t)f.gcount() != n * sizeof(T)) {
        err = path + ": expected " + std::to_string(n * sizeof(T)) + " bytes, got " + std::to_string(f.gcount());
        return false;
    }
    return true;
}

template <class T>
bool write_all(const std::string &path, const T *src, size_t n) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char *>(src), (std::streamsize)(n * sizeof(T)));
    return (bool)f;
}

}  // namespace

bool read_metaint(const std::string &dir, MetaInt &m, std::string &err) {
    int32_t v[37];
    if (!read_all(dir + "/bin/metaint.bin", v, 37, err)) return false;
    int k = 0;
    int *scal[] = {&m.gpu, &m.surf, &m.pml_z, &m.pml_x, &m.accu_save, &m.seismo_save, &m.fwinv, &m.rtf_meas_true,
                   &m.nt, &m.nz, &m.nx, &m.snap_t1, &m.snap_t2, &m.snap_z1, &m.snap_z2, &m.snap_x1, &m.snap_x2,
                   &m.snap_dt, &m.snap_dz, &m.snap_dx};
    for (int *p : scal) *p = v[k++];
    for (int i = 0; i < 8; ++i) m.taper[i] = v[k++];
    int *tail[] = {&m.nsrc, &m.nrec, &m.nshot, &m.stf_type, &m.rtf_type, &m.fdorder, &m.fpad,
                   &m.mat_save_interval, &m.mat_grid};
    for (int *p : tail) *p = v[k++];
    return true;
}

bool read_intarray(const std::string &dir, const MetaInt &m, IntArrays &a, std::string &err) {
    const size_t n = 8 + 3 * (size_t)m.nsrc + 2 * (size_t)m.nrec;
    std::vector<int32_t> v(n);
    if (!read_all(dir + "/bin/intarray.bin", v.data(), n, err)) return false;
    for (int i = 0; i < 4; ++i) a.npml[i] = v[i];
    for (int i = 0; i < 4; ++i) a.isurf[i] = v[4 + i];
    size_t k = 8;
    auto take = [&](std::vector<int32_t> &dst, int cnt) {
        dst.assign(v.begin() + k, v.begin() + k + cnt);
        k += cnt;
    };
    take(a.z_src, m.nsrc);
    take(a.x_src, m.nsrc);
    take(a.fire, m.nsrc);
    take(a.z_rec, m.nrec);
    take(a.x_rec, m.nrec);
    return true;
}

bool read_metafloat(const std::string &dir, MetaFloat &f, std::string &err) {
    double v[11];
    if (!read_all(dir + "/bin/metafloat.bin", v, 11, err)) return false;
    f = MetaFloat{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10]};
    return true;
}

bool read_material(const std::string &dir, int nz, int nx, std::vector<double> &lam, std::vector<double> &mu,
                   std::vector<double> &rho, std::string &err) {
    const size_t n = (size_t)nz * nx;
    std::vector<double> v(3 * n);
    if (!read_all(dir + "/bin/mat.bin", v.data(), 3 * n, err)) return false;
    lam.assign(v.begin(), v.begin() + n);
    mu.assign(v.begin() + n, v.begin() + 2 * n);
    rho.assign(v.begin() + 2 * n, v.end());
    return true;
}

bool read_seismo(const std::string &dir, int ishot, int nrec, int nt, double *uz, double *ux, std::string &err) {
    const std::string p = dir + "/io/shot" + std::to_string(ishot) + "_rtf_";
    return read_all(p + "uz.bin", uz, (size_t)nrec * nt, err) && read_all(p + "ux.bin", ux, (size_t)nrec * nt, err);
}

bool write_seismo(const std::string &dir, int ishot, int nrec, int nt, const double *uz, const double *ux) {
    const std::string p = dir + "/bin/shot" + std::to_string(ishot) + "_rtf_";
    return write_all(p + "uz.bin", uz, (size_t)nrec * nt) && write_all(p + "ux.bin", ux, (size_t)nrec * nt);
}

bool write_accu(const std::string &dir, int ishot, size_t n, const double *const fields[5]) {
    static const char *names[5] = {"vz", "vx", "szz", "szx", "sxx"};
    bool ok = true;
    for (int k = 0; k < 5; ++k)
        ok = write_all(dir + "/bin/shot" + std::to_string(ishot) + "_" + names[k] + ".bin", fields[k], n) && ok;
    return ok;
}

bool write_mat(const std::string &dir, int iter, size_t n, const double *lam, const double *mu, const double *rho) {
    std::ofstream f(dir + "/bin/iter" + std::to_string(iter) + "_mat.bin", std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char *>(lam), (std::streamsize)(n * 8));
    f.write(reinterpret_cast<const char *>(mu), (std::streamsize)(n * 8));
    f.write(reinterpret_cast<const char *>(rho), (std::streamsize)(n * 8));
    return (bool)f;
}

}  // namespace psvio
