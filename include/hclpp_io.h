/* hclpp_io.h -- output and restart files of the hydro state in the reference's HDF5 layout
 * (libhclpp_io.so, host code over the HDF5 C library; no GPU code).
 *
 * The reference writes its outputs through PDI's decl_hdf5 plugin with the layout declared in
 * src/io/config.yaml.cpp:53-239 and reads restarts the same way (src/io/io.cpp:185-431).  One
 * file per output, <directory>/<prefix>_<output_id, 8 digits>.h5 (io.cpp:62-74), holding at
 * its root:
 *   output_id, iter_output_id, time_output_id, iter     int scalars
 *   current_time, gamma                                 double scalars
 *   x0, x1, x2        node coordinates with ghosts, nx[d] + 2 ng[d] + 1 values each
 *   x0_ng, x1_ng, x2_ng  the interior nodes, nx[d] + 1 values (from index ng[d])
 *   rho, ux0 (ux1 when ndim > 1, ux2 when ndim > 2), P, E, T
 *                     interior cells, dataset shape [nx2][nx1][nx0] (x fastest)
 *   fx0 ... fx<nfx-1> passive scalars, same shape
 *   string attributes git_build_string, git_branch, compile_date, compile_time on the file.
 * Arrays passed in and out are HOST arrays in the solver's ghost-inclusive LayoutLeft layout
 * (index i + nwg0 (j + nwg1 k), one plane per component); only their interior is written or
 * filled, as the reference's memory_selection does.  Every function returns HCLPP_IO_OK or an
 * error code, with the message in hclpp_io_last_error(); nothing is thrown across the ABI.
 */
#ifndef HCLPP_IO_H
#define HCLPP_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum { HCLPP_IO_OK = 0, HCLPP_IO_ERR_ARG = 1, HCLPP_IO_ERR_HDF5 = 2, HCLPP_IO_ERR_LAYOUT = 3 };

typedef struct {
    int ndim;              /* 1, 2 or 3 */
    int nfx;               /* passive scalars (at most 10 datasets fx0..fx9, config.yaml.cpp:155-165) */
    int n[3];              /* interior cells per dim (1 on inactive dims): Nx_glob_ng */
    int ng;                /* ghost layers on active dims: Nghost = ng for d < ndim, else 0 */
    const double* x[3];    /* node coordinates with ghosts, n[d] + 2 Nghost[d] + 1 values (host) */
} hclpp_io_grid;

typedef struct {
    int output_id;
    int iter_output_id;
    int time_output_id;
    int iter;
    double current_time;
    double gamma;
} hclpp_io_header;

const char* hclpp_io_last_error(void);

/* The output file name of main.cpp / io.cpp:62-74: prefix + "_" + output_id (8 digits, zero
 * filled) + ".h5", written into out (size bytes); fails when it does not fit. */
int hclpp_io_output_filename(const char* prefix, int output_id, char* out, int size);

/* write_pdi (io/io.cpp:185-324): create or replace <directory>/<name> and write the header,
 * the node coordinates and the interior of rho, u (ndim planes), P, E, T and fx (nfx planes;
 * NULL when nfx = 0). */
int hclpp_io_write(const hclpp_io_grid* grid, const char* directory, const char* prefix, const hclpp_io_header* header,
                   const double* rho, const double* u, const double* P, const double* E, const double* T, const double* fx);

/* read_pdi (io/io.cpp:326-431): read the header (its gamma field is set from the file too),
 * the node coordinates (x0, x1, x2: the extents must match the grid's, io.cpp:364-366) and the
 * interior of rho, u, P and fx (nfx planes; NULL when nfx = 0) into the given host arrays; the
 * ghost entries of those arrays are left as they were.  x_out[d] receive n[d] + 2 Nghost[d] + 1
 * values each (NULL to skip). */
int hclpp_io_read(const hclpp_io_grid* grid, const char* filename, hclpp_io_header* header, double* rho, double* u, double* P,
                  double* fx, double* x_out0, double* x_out1, double* x_out2);

/* One interior dataset of an output file (rho, ux0, P, E, T, fx0, ...), into the interior of a
 * ghost-inclusive host plane (for post-processing and checks; ghost entries are left as they were). */
int hclpp_io_read_field(const hclpp_io_grid* grid, const char* filename, const char* dataset, double* plane);

#ifdef __cplusplus
}
#endif
#endif
