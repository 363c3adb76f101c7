#include "Grid.h"

// Grid::Grid (reference src/Grid.cpp:4-24): nx_ interior points on
// [start,end] plus a halo cell on each side.
Grid::Grid(size_t nx_, size_t ny_, Coord start_, Coord end_, size_t halo)
{
    if (end_[1] < start_[1]) ExitOnError("Grid x in [a,b] has b<a!");
    if (end_[2] < start_[2]) ExitOnError("Grid y in [a,b] has b<a!");
    if (nx_ <= 1 || ny_ <= 1) ExitOnError("Grid must have at least 2 LP in each Dimension.");
    nx = nx_ + 2 * halo;
    ny = ny_ + 2 * halo;
    nxy = nx * ny;
    dx = (end_[1] - start_[1]) / (nx - 1.0 - 2.0 * halo);
    dy = (end_[2] - start_[2]) / (ny - 1.0 - 2.0 * halo);
    dt = m_cfl * std::min(dx, dy);
    startx = start_[1] - dx;
    starty = start_[2] - dy;
    endx = end_[1] + dx;
    endy = end_[2] + dy;
}
void Grid::SetCFL(double cfl)
{
    m_cfl = cfl;
    dt = m_cfl * std::min(dx, dy);
}
