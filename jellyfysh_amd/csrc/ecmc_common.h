// ecmc_common.h -- host-side helpers shared by the HIP translation units: error reporting
// and the set-up tables (cell grid, Ewald Fourier factors).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

// Records `msg` as the calling thread's last error and returns `status`.
int ecmc_fail(int status, const char *msg);

namespace ecmc {

// CuboidPeriodicCells (activator/internal_state/cell_occupancy/cells/cuboid_cells.py:54-146,
// cuboid_periodic_cells.py:74-100): cells indexed with the first dimension fastest, float-exact
// cell bounds (the smallest/largest doubles that int(x / side) maps into the cell), and for
// each cell the nearby cells in the order itertools.product yields them, wrapped and with
// repeats dropped.
struct CellGrid {
    int dim = 0, count = 0, nb_stride = 0;
    int cps[3] = {1, 1, 1};
    long cum[3] = {0, 0, 0};
    double side[3] = {1, 1, 1};
    std::vector<double> cmin, cmax;     // count x dim
    std::vector<int32_t> nearby;        // count x nb_stride
    std::vector<int32_t> nearby_n;      // count
    int position_to_cell(const double *p) const;
};
bool build_cell_grid(int dim, double system_length, const int32_t *cells_per_side, int neighbor_layers, CellGrid &g,
                     std::string &err);

// Fourier factors of the merged-image Coulomb derivative, index (i * (fc+1) + j) * (fc+1) + k,
// evaluated in the order of merged_image_coulomb_potential.c:94-109 so the table is the
// reference's bit for bit.
void build_ewald_table(int fourier_cutoff, double alpha, double system_length, std::vector<double> &table);

}  // namespace ecmc
