// io_hdf5.cpp -- libhclpp_io.so: output and restart files in the reference's HDF5 layout
// (include/hclpp_io.h), on the HDF5 C library.  Host code only: the caller copies the device
// state to the host first, as the reference's write_pdi does through host views.
//
// Layout: src/io/config.yaml.cpp:53-239 (the decl_hdf5 datasets PDI writes and reads); write and
// read sequences: src/io/io.cpp:185-324 (write_pdi) and :326-431 (read_pdi).  The interior of a
// ghost-inclusive array is selected with a memory hyperslab starting at Nghost, as the
// reference's memory_selection does, so no staging copy is made.
#include <hdf5.h>

#include <array>
#include <cstdio>
#include <string>

#include "hclpp_io.h"

namespace {

thread_local std::string g_error;

int fail(int code, std::string msg)
{
    g_error = std::move(msg);
    return code;
}

// Owner of an HDF5 identifier, closed with the matching H5*close.
class Hid
{
    hid_t m_id;
    herr_t (*m_close)(hid_t);

public:
    Hid(hid_t id, herr_t (*close)(hid_t)) : m_id(id), m_close(close) {}
    Hid(Hid const&) = delete;
    Hid& operator=(Hid const&) = delete;
    ~Hid()
    {
        if (m_id >= 0) m_close(m_id);
    }
    hid_t operator*() const { return m_id; }
    bool ok() const { return m_id >= 0; }
};

struct Extents {
    int nd;
    std::array<int, 3> n, ngh, nwg;
};

int check_grid(const hclpp_io_grid* g, Extents& e)
{
    if (!g) return fail(HCLPP_IO_ERR_ARG, "null grid");
    if (g->ndim < 1 || g->ndim > 3) return fail(HCLPP_IO_ERR_ARG, "ndim must be 1, 2 or 3");
    if (g->nfx < 0 || g->nfx > 10) return fail(HCLPP_IO_ERR_ARG, "nfx must be within 0..10 (datasets fx0..fx9)");
    if (g->ng < 0) return fail(HCLPP_IO_ERR_ARG, "ng < 0");
    e.nd = g->ndim;
    for (int d = 0; d < 3; ++d) {
        if (g->n[d] < 1 || (d >= g->ndim && g->n[d] != 1)) return fail(HCLPP_IO_ERR_ARG, "invalid interior extents");
        e.n[d] = g->n[d];
        e.ngh[d] = d < g->ndim ? g->ng : 0;
        e.nwg[d] = e.n[d] + 2 * e.ngh[d];
    }
    return HCLPP_IO_OK;
}

std::string name_of(const char* prefix, int output_id)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "_%08d.h5", output_id);
    return std::string(prefix) + buf;
}

int write_scalar(hid_t file, const char* name, hid_t type, const void* v)
{
    Hid const space(H5Screate(H5S_SCALAR), H5Sclose);
    Hid const set(H5Dcreate2(file, name, type, *space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!set.ok() || H5Dwrite(*set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
        return fail(HCLPP_IO_ERR_HDF5, std::string("writing ") + name + " failed");
    return HCLPP_IO_OK;
}

int write_1d(hid_t file, const char* name, const double* v, hsize_t n, hsize_t start)
{
    hsize_t const dim = n;
    Hid const fspace(H5Screate_simple(1, &dim, nullptr), H5Sclose);
    Hid const set(H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, *fspace, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!set.ok() || H5Dwrite(*set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v + start) < 0)
        return fail(HCLPP_IO_ERR_HDF5, std::string("writing ") + name + " failed");
    return HCLPP_IO_OK;
}

// memory space of one ghost-inclusive plane with its interior selected (memory_selection)
hid_t interior_memspace(const Extents& e)
{
    hsize_t const mdims[3] = {(hsize_t)e.nwg[2], (hsize_t)e.nwg[1], (hsize_t)e.nwg[0]};
    hsize_t const start[3] = {(hsize_t)e.ngh[2], (hsize_t)e.ngh[1], (hsize_t)e.ngh[0]};
    hsize_t const count[3] = {(hsize_t)e.n[2], (hsize_t)e.n[1], (hsize_t)e.n[0]};
    hid_t const m = H5Screate_simple(3, mdims, nullptr);
    if (m >= 0 && H5Sselect_hyperslab(m, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0) {
        H5Sclose(m);
        return -1;
    }
    return m;
}

int write_interior(hid_t file, const char* name, const double* plane, const Extents& e)
{
    hsize_t const fdims[3] = {(hsize_t)e.n[2], (hsize_t)e.n[1], (hsize_t)e.n[0]};
    Hid const fspace(H5Screate_simple(3, fdims, nullptr), H5Sclose);
    Hid const mspace(interior_memspace(e), H5Sclose);
    Hid const set(H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, *fspace, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!mspace.ok() || !set.ok() || H5Dwrite(*set, H5T_NATIVE_DOUBLE, *mspace, H5S_ALL, H5P_DEFAULT, plane) < 0)
        return fail(HCLPP_IO_ERR_HDF5, std::string("writing ") + name + " failed");
    return HCLPP_IO_OK;
}

int write_string_attribute(hid_t file, const char* name, const std::string& value)
{
    Hid const space(H5Screate(H5S_SCALAR), H5Sclose);
    Hid const type(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!type.ok() || H5Tset_size(*type, value.empty() ? 1 : value.size()) < 0 || H5Tset_cset(*type, H5T_CSET_UTF8) < 0)
        return fail(HCLPP_IO_ERR_HDF5, std::string("string type for ") + name);
    Hid const attr(H5Acreate2(file, name, *type, *space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok() || H5Awrite(*attr, *type, value.c_str()) < 0) return fail(HCLPP_IO_ERR_HDF5, std::string("writing attribute ") + name);
    return HCLPP_IO_OK;
}

// dataset dims, or a layout error when absent or of another rank
int dims_of(hid_t file, const char* name, int rank, hsize_t* dims)
{
    if (H5Lexists(file, name, H5P_DEFAULT) <= 0) return fail(HCLPP_IO_ERR_LAYOUT, std::string("dataset ") + name + " not found");
    Hid const set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    Hid const space(H5Dget_space(*set), H5Sclose);
    if (!set.ok() || !space.ok()) return fail(HCLPP_IO_ERR_HDF5, std::string("opening ") + name);
    int const r = H5Sget_simple_extent_ndims(*space);
    if (r != rank) return fail(HCLPP_IO_ERR_LAYOUT, std::string("dataset ") + name + " has rank " + std::to_string(r));
    if (rank > 0) H5Sget_simple_extent_dims(*space, dims, nullptr);
    return HCLPP_IO_OK;
}

int read_scalar(hid_t file, const char* name, hid_t type, void* v)
{
    int rc = dims_of(file, name, 0, nullptr);
    if (rc) return rc;
    Hid const set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (H5Dread(*set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) return fail(HCLPP_IO_ERR_HDF5, std::string("reading ") + name);
    return HCLPP_IO_OK;
}

int read_interior(hid_t file, const char* name, double* plane, const Extents& e)
{
    hsize_t dims[3];
    int rc = dims_of(file, name, 3, dims);
    if (rc) return rc;
    if (dims[0] != (hsize_t)e.n[2] || dims[1] != (hsize_t)e.n[1] || dims[2] != (hsize_t)e.n[0])
        return fail(HCLPP_IO_ERR_LAYOUT, std::string("dataset ") + name + " does not have the grid's interior extents");
    Hid const set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    Hid const mspace(interior_memspace(e), H5Sclose);
    if (!mspace.ok() || H5Dread(*set, H5T_NATIVE_DOUBLE, *mspace, H5S_ALL, H5P_DEFAULT, plane) < 0)
        return fail(HCLPP_IO_ERR_HDF5, std::string("reading ") + name);
    return HCLPP_IO_OK;
}

}  // namespace

extern "C" {

const char* hclpp_io_last_error(void) { return g_error.c_str(); }

int hclpp_io_output_filename(const char* prefix, int output_id, char* out, int size)
{
    if (!prefix || !out || size <= 0 || output_id < 0) return fail(HCLPP_IO_ERR_ARG, "invalid argument");
    std::string const s = name_of(prefix, output_id);
    if ((int)s.size() + 1 > size) return fail(HCLPP_IO_ERR_ARG, "buffer too small for the file name");
    std::snprintf(out, (size_t)size, "%s", s.c_str());
    return HCLPP_IO_OK;
}

int hclpp_io_write(const hclpp_io_grid* grid, const char* directory, const char* prefix, const hclpp_io_header* h, const double* rho,
                   const double* u, const double* P, const double* E, const double* T, const double* fx)
{
    Extents e;
    int rc = check_grid(grid, e);
    if (rc) return rc;
    if (!directory || !prefix || !h || !rho || !u || !P || !E || !T || (grid->nfx > 0 && !fx)) return fail(HCLPP_IO_ERR_ARG, "null argument");
    for (int d = 0; d < 3; ++d)
        if (!grid->x[d]) return fail(HCLPP_IO_ERR_ARG, "null node coordinates");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    std::string const path = std::string(directory) + "/" + name_of(prefix, h->output_id);
    // collision_policy: replace (config.yaml.cpp:56)
    Hid const file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(HCLPP_IO_ERR_HDF5, "cannot create " + path);
    hid_t const f = *file;
    // write_replicated_data (config.yaml.cpp:53-85)
    if ((rc = write_scalar(f, "output_id", H5T_NATIVE_INT, &h->output_id))) return rc;
    if ((rc = write_scalar(f, "iter_output_id", H5T_NATIVE_INT, &h->iter_output_id))) return rc;
    if ((rc = write_scalar(f, "time_output_id", H5T_NATIVE_INT, &h->time_output_id))) return rc;
    if ((rc = write_scalar(f, "iter", H5T_NATIVE_INT, &h->iter))) return rc;
    if ((rc = write_scalar(f, "current_time", H5T_NATIVE_DOUBLE, &h->current_time))) return rc;
    if ((rc = write_scalar(f, "gamma", H5T_NATIVE_DOUBLE, &h->gamma))) return rc;
    char const* const xn[3] = {"x0", "x1", "x2"};
    char const* const xng[3] = {"x0_ng", "x1_ng", "x2_ng"};
    for (int d = 0; d < 3; ++d) {
        if ((rc = write_1d(f, xn[d], grid->x[d], (hsize_t)e.nwg[d] + 1, 0))) return rc;
        if ((rc = write_1d(f, xng[d], grid->x[d], (hsize_t)e.n[d] + 1, (hsize_t)e.ngh[d]))) return rc;
    }
    // write_distributed_data (config.yaml.cpp:86-150)
    size_t const plane = (size_t)e.nwg[0] * e.nwg[1] * e.nwg[2];
    char const* const un[3] = {"ux0", "ux1", "ux2"};
    for (int d = 0; d < e.nd; ++d)
        if ((rc = write_interior(f, un[d], u + plane * d, e))) return rc;
    if ((rc = write_interior(f, "rho", rho, e))) return rc;
    if ((rc = write_interior(f, "P", P, e))) return rc;
    if ((rc = write_interior(f, "E", E, e))) return rc;
    if ((rc = write_interior(f, "T", T, e))) return rc;
    // write_fx (config.yaml.cpp:151-172)
    for (int ifx = 0; ifx < grid->nfx; ++ifx) {
        std::string const name = "fx" + std::to_string(ifx);
        if ((rc = write_interior(f, name.c_str(), fx + plane * ifx, e))) return rc;
    }
    // the build attributes write_pdi adds (io.cpp:316-322)
    if ((rc = write_string_attribute(f, "git_build_string", "heraclespp_amd hydro step for MI355X (libhclpp_hip)"))) return rc;
    if ((rc = write_string_attribute(f, "git_branch", ""))) return rc;
    if ((rc = write_string_attribute(f, "compile_date", __DATE__))) return rc;
    if ((rc = write_string_attribute(f, "compile_time", __TIME__))) return rc;
    return HCLPP_IO_OK;
}

int hclpp_io_read(const hclpp_io_grid* grid, const char* filename, hclpp_io_header* h, double* rho, double* u, double* P, double* fx,
                  double* x_out0, double* x_out1, double* x_out2)
{
    Extents e;
    int rc = check_grid(grid, e);
    if (rc) return rc;
    if (!filename || !h || !rho || !u || !P || (grid->nfx > 0 && !fx)) return fail(HCLPP_IO_ERR_ARG, "null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Hid const file(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(HCLPP_IO_ERR_HDF5, std::string("cannot open ") + filename);
    hid_t const f = *file;
    // check_extent_dset of x0, x1, x2 (io.cpp:364-366)
    char const* const xn[3] = {"x0", "x1", "x2"};
    double* const xo[3] = {x_out0, x_out1, x_out2};
    for (int d = 0; d < 3; ++d) {
        hsize_t n = 0;
        if ((rc = dims_of(f, xn[d], 1, &n))) return rc;
        if (n != (hsize_t)e.nwg[d] + 1) return fail(HCLPP_IO_ERR_LAYOUT, std::string("extent of ") + xn[d] + " does not match the grid");
        if (xo[d]) {
            Hid const set(H5Dopen2(f, xn[d], H5P_DEFAULT), H5Dclose);
            if (H5Dread(*set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, xo[d]) < 0)
                return fail(HCLPP_IO_ERR_HDF5, std::string("reading ") + xn[d]);
        }
    }
    // read_file (config.yaml.cpp:174-222)
    if ((rc = read_scalar(f, "output_id", H5T_NATIVE_INT, &h->output_id))) return rc;
    if ((rc = read_scalar(f, "iter_output_id", H5T_NATIVE_INT, &h->iter_output_id))) return rc;
    if ((rc = read_scalar(f, "time_output_id", H5T_NATIVE_INT, &h->time_output_id))) return rc;
    if ((rc = read_scalar(f, "iter", H5T_NATIVE_INT, &h->iter))) return rc;
    if ((rc = read_scalar(f, "current_time", H5T_NATIVE_DOUBLE, &h->current_time))) return rc;
    if ((rc = read_scalar(f, "gamma", H5T_NATIVE_DOUBLE, &h->gamma))) return rc;
    size_t const plane = (size_t)e.nwg[0] * e.nwg[1] * e.nwg[2];
    if ((rc = read_interior(f, "rho", rho, e))) return rc;
    char const* const un[3] = {"ux0", "ux1", "ux2"};
    for (int d = 0; d < e.nd; ++d)
        if ((rc = read_interior(f, un[d], u + plane * d, e))) return rc;
    if ((rc = read_interior(f, "P", P, e))) return rc;
    // read_fx (config.yaml.cpp:223-233)
    for (int ifx = 0; ifx < grid->nfx; ++ifx) {
        std::string const name = "fx" + std::to_string(ifx);
        if ((rc = read_interior(f, name.c_str(), fx + plane * ifx, e))) return rc;
    }
    return HCLPP_IO_OK;
}

int hclpp_io_read_field(const hclpp_io_grid* grid, const char* filename, const char* dataset, double* plane)
{
    Extents e;
    int const rc = check_grid(grid, e);
    if (rc) return rc;
    if (!filename || !dataset || !plane) return fail(HCLPP_IO_ERR_ARG, "null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Hid const file(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(HCLPP_IO_ERR_HDF5, std::string("cannot open ") + filename);
    return read_interior(*file, dataset, plane, e);
}

}  // extern "C"
