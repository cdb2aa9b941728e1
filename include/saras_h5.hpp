// saras_h5.hpp -- HDF5 solution and restart files of a saras_amd::lsRK3 run
// in the reference's layout, for callers that link HDF5 as the reference
// does (header-only; needs hdf5.h and -lhdf5).
//
//   saras_amd::writer  stands for writer (lib/io/writer.cc:331-497):
//                      writeSolution(time) -> <dir>/Soln_TTTT.TTTT.h5,
//                      writeRestart(time)  -> <dir>/restartFile.h5
//   saras_amd::reader  stands for reader (lib/io/reader.cc:171-437):
//                      readRestart() loads <dir>/restartFile.h5 into the
//                      handle and returns its time
//
// Layout (writer.cc:60-200): a scalar dataset "Time"; in solution files the
// 1-D datasets "X", "Y", "Z" with the core cell-centre coordinates; one
// (n_x, n_y, n_z) dataset of the core per field, C order, named "Vx", "Vy",
// "Vz", "P" and, with a scalar, "T".  Planar (dim 2) runs follow the PLANAR
// build (writer.cc:80-160, 361-367; hydro_d2.cc:137-139, scalar_d2.cc:137-140):
// (n_x, n_z) datasets of "Vx", "Vz", "P" and "T", coordinates "X" and "Z"
// only; their reader expects rank 2 (reader.cc:377-437) and leaves Vy at
// zero.  One process holds the whole domain, so
// the files are written with serial HDF5 (the reference's MPI-IO property
// lists select the same bytes).  The reader's consistency check (fileCheck,
// reader.cc:377-437: rank 3, extents equal to the grid's) throws
// saras_amd::error instead of exit(0).
#pragma once
#include <hdf5.h>

#include <cstdio>
#include <string>
#include <vector>

#include "saras_dropin.hpp"

namespace saras_amd {

namespace h5detail {

inline const char *const kFieldNames[5] = {"Vx", "Vy", "Vz", "P", "T"};

// full (pad) arrays <-> contiguous core blocks; planar arrays have no y pad
struct Layout {
    int n[3], pad;
    bool planar;
    int ypad() const { return planar ? 0 : pad; }
    size_t full() const { return (size_t)(n[0] + 2 * pad) * (n[1] + 2 * ypad()) * (n[2] + 2 * pad); }
    size_t core() const { return (size_t)n[0] * n[1] * n[2]; }
    size_t at(int i, int j, int k) const {
        return ((size_t)(i + pad) * (n[1] + 2 * ypad()) + (j + ypad())) * (n[2] + 2 * pad) + (k + pad);
    }
    int rank() const { return planar ? 2 : 3; }
    // dataset extents: (n_x, n_z) planar, (n_x, n_y, n_z) otherwise
    int dims(hsize_t *d) const {
        d[0] = (hsize_t)n[0];
        d[1] = (hsize_t)(planar ? n[2] : n[1]);
        d[2] = (hsize_t)n[2];
        return rank();
    }
    // the fields a run writes: Vx, Vy, Vz, P, T by index; planar drops Vy
    bool writes(int q, bool scalar) const { return !(planar && q == 1) && (q < 4 || scalar); }
    void to_core(const std::vector<double> &f, std::vector<double> &c) const {
        c.resize(core());
        size_t q = 0;
        for (int i = 0; i < n[0]; ++i)
            for (int j = 0; j < n[1]; ++j)
                for (int k = 0; k < n[2]; ++k) c[q++] = f[at(i, j, k)];
    }
    void from_core(const std::vector<double> &c, std::vector<double> &f) const {
        f.assign(full(), 0.0);
        size_t q = 0;
        for (int i = 0; i < n[0]; ++i)
            for (int j = 0; j < n[1]; ++j)
                for (int k = 0; k < n[2]; ++k) f[at(i, j, k)] = c[q++];
    }
};

inline Layout layout_of(const saras_grid_desc &g) {
    if (g.dim != 3 && !(g.dim == 2 && g.n[1] == 1))
        throw error(SARAS_BAD_SHAPE, "HDF5 files: 3-D grids, or planar grids with n_y = 1");
    Layout L;
    for (int a = 0; a < 3; ++a) L.n[a] = g.n[a];
    L.pad = g.pad;
    L.planar = g.dim == 2;
    return L;
}

inline void write_1d(hid_t file, const char *name, const std::vector<double> &v) {
    const hsize_t d = v.size();
    const hid_t sp = H5Screate_simple(1, &d, nullptr);
    const hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t st = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    if (st < 0) throw error(SARAS_BAD_ARG, std::string("HDF5 write of ") + name + " failed");
}

inline void write_time(hid_t file, double time) {
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t ds = H5Dcreate2(file, "Time", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t st = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_DOUBLE, sp, sp, H5P_DEFAULT, &time);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    if (st < 0) throw error(SARAS_BAD_ARG, "HDF5 write of Time failed");
}

inline void write_field(hid_t file, const char *name, const Layout &L, const std::vector<double> &core) {
    hsize_t d[3];
    const hid_t sp = H5Screate_simple(L.dims(d), d, nullptr);
    const hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t st = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, core.data());
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    if (st < 0) throw error(SARAS_BAD_ARG, std::string("HDF5 write of ") + name + " failed");
}

}  // namespace h5detail

class writer {
  public:
    writer(lsRK3 &solver, const saras_grid_desc &grid, bool scalar, std::string dir = "output")
        : ts_(solver), g_(grid), L_(h5detail::layout_of(grid)), scalar_(scalar), dir_(std::move(dir)) {}

    // writer::writeSolution (writer.cc:331-405): coordinates, time, fields
    std::string writeSolution(double time) {
        char name[64];
        std::snprintf(name, sizeof(name), "Soln_%09.4f.h5", time);
        const std::string path = dir_ + "/" + name;
        const hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (f < 0) throw error(SARAS_BAD_ARG, "cannot create " + path);
        try {
            static const char *const axes[3] = {"X", "Y", "Z"};
            for (int a = 0; a < 3; ++a) {
                if (L_.planar && a == 1) continue;   // writer.cc:361-366
                std::vector<double> c(L_.n[a] + 2), xi(L_.n[a] + 2), x1(L_.n[a] + 2), x2(L_.n[a] + 2), x3(L_.n[a] + 2);
                check(saras_level_metrics(&g_, 0, a, c.data(), xi.data(), x1.data(), x2.data(), x3.data()),
                      "saras_level_metrics");
                h5detail::write_1d(f, axes[a], std::vector<double>(c.begin() + 1, c.end() - 1));
            }
            h5detail::write_time(f, time);
            write_fields(f);
        } catch (...) {
            H5Fclose(f);
            throw;
        }
        H5Fclose(f);
        return path;
    }

    // writer::writeRestart (writer.cc:417-478): time and fields
    std::string writeRestart(double time) {
        const std::string path = dir_ + "/restartFile.h5";
        const hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (f < 0) throw error(SARAS_BAD_ARG, "cannot create " + path);
        try {
            h5detail::write_time(f, time);
            write_fields(f);
        } catch (...) {
            H5Fclose(f);
            throw;
        }
        H5Fclose(f);
        return path;
    }

  private:
    void write_fields(hid_t f) {
        std::vector<std::vector<double>> full(5, std::vector<double>(L_.full(), 0.0));
        ts_.getFields(full[0].data(), full[1].data(), full[2].data(), full[3].data(),
                      scalar_ ? full[4].data() : nullptr);
        std::vector<double> core;
        for (int q = 0; q < 5; ++q) {
            if (!L_.writes(q, scalar_)) continue;
            L_.to_core(full[q], core);
            h5detail::write_field(f, h5detail::kFieldNames[q], L_, core);
        }
    }

    lsRK3 &ts_;
    saras_grid_desc g_;
    h5detail::Layout L_;
    bool scalar_;
    std::string dir_;
};

class reader {
  public:
    reader(lsRK3 &solver, const saras_grid_desc &grid, bool scalar, std::string dir = "output")
        : ts_(solver), L_(h5detail::layout_of(grid)), scalar_(scalar), dir_(std::move(dir)) {}

    // reader::readRestart (reader.cc:171-260): fileCheck, time, fields; the
    // fields are loaded into the handle, whose boundary conditions apply
    double readRestart() {
        const std::string path = dir_ + "/restartFile.h5";
        hid_t f;
        H5E_BEGIN_TRY { f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT); }
        H5E_END_TRY;
        if (f < 0) throw error(SARAS_BAD_ARG, "could not open restart file " + path);
        double time = 0.0;
        std::vector<std::vector<double>> full(5, std::vector<double>(L_.full(), 0.0));
        try {
            fileCheck(f);
            const hid_t ds = H5Dopen2(f, "Time", H5P_DEFAULT);
            if (ds < 0 || H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &time) < 0) {
                if (ds >= 0) H5Dclose(ds);
                throw error(SARAS_BAD_ARG, "HDF5 read of Time failed");
            }
            H5Dclose(ds);
            std::vector<double> core(L_.core());
            for (int q = 0; q < 5; ++q) {
                if (!L_.writes(q, scalar_)) continue;
                const hid_t d = H5Dopen2(f, h5detail::kFieldNames[q], H5P_DEFAULT);
                if (d < 0 || H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, core.data()) < 0) {
                    if (d >= 0) H5Dclose(d);
                    throw error(SARAS_BAD_ARG, std::string("HDF5 read of ") + h5detail::kFieldNames[q] + " failed");
                }
                H5Dclose(d);
                L_.from_core(core, full[q]);
            }
        } catch (...) {
            H5Fclose(f);
            throw;
        }
        H5Fclose(f);
        ts_.setFields(full[0].data(), full[1].data(), full[2].data(), full[3].data(),
                      scalar_ ? full[4].data() : nullptr);
        return time;
    }

  private:
    // reader::fileCheck (reader.cc:377-437): the pressure dataset has the
    // run's rank (3, or 2 for planar runs) and the grid's extents
    void fileCheck(hid_t f) const {
        hid_t d;
        H5E_BEGIN_TRY { d = H5Dopen2(f, "P", H5P_DEFAULT); }
        H5E_END_TRY;
        if (d < 0) throw error(SARAS_BAD_SHAPE, "restart file holds no P dataset");
        const hid_t sp = H5Dget_space(d);
        const int nd = H5Sget_simple_extent_ndims(sp);
        hsize_t want[3], dims[3] = {0, 0, 0};
        const int r = L_.dims(want);
        bool ok = nd == r && H5Sget_simple_extent_dims(sp, dims, nullptr) == r;
        for (int a = 0; ok && a < r; ++a) ok = dims[a] == want[a];
        H5Sclose(sp);
        H5Dclose(d);
        if (nd != r) throw error(SARAS_BAD_SHAPE, "Dimensionality of restart file conflicts with solver parameters");
        if (!ok) throw error(SARAS_BAD_SHAPE, "Array limits in restart file conflict with solver parameters");
    }

    lsRK3 &ts_;
    h5detail::Layout L_;
    bool scalar_;
    std::string dir_;
};

}  // namespace saras_amd
