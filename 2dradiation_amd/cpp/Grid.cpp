#include "Grid.h"
#include <fstream>

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

// Grid::WriteFrametoCsv (reference src/Grid.cpp:109-136): one frame as text,
// file "<directory><name><Format(time, 3)>t <nx>x <ny>y.csv" with a '#'
// header (grid extent and time) and one line "x,y,r,g,b,a" per cell in
// storage order, numbers in the stream's default format.
void Grid::WriteFrametoCsv(float time, const DoubleBuffer &r, const DoubleBuffer &g, const DoubleBuffer &b,
                           const DoubleBuffer &a, std::string directory, std::string name)
{
    CreateDirectory(directory);
    if (name.empty()) name = "data";
    std::ofstream out(directory + name + Format(time, 3) + "t " + std::to_string(nx) + "x " + std::to_string(ny) +
                      "y" + ".csv");
    out << "#nx=" << nx << "\n";
    out << "#ny=" << ny << "\n";
    out << "#startx=" << startx << "\n";
    out << "#starty=" << starty << "\n";
    out << "#endx=" << endx << "\n";
    out << "#endy=" << endy << "\n";
    out << "#t=" << time << "\n";
    out << "#x,y,r,g,b,a\n";
    for (size_t j = 0; j < ny; j++)
        for (size_t i = 0; i < nx; i++) {
            const size_t ij = Index(i, j);
            out << x(i) << "," << y(j) << ",";
            out << r[ij] << "," << g[ij] << "," << b[ij] << "," << a[ij] << "\n";
        }
}

// Grid::Print (reference src/Grid.cpp:140-153).
void Grid::Print() const
{
    std::cout << "Grid: nx=" << nx << "\n";
    std::cout << "Grid: ny=" << ny << "\n";
    std::cout << "Grid: nxy=" << nxy << "\n";
    std::cout << "Grid: dx=" << dx << "\n";
    std::cout << "Grid: dy=" << dy << "\n";
    std::cout << "Grid: dt=" << dt << "\n";
    std::cout << "Grid: startx=" << startx << "\n";
    std::cout << "Grid: starty=" << starty << "\n";
    std::cout << "Grid: endx=" << endx << "\n";
    std::cout << "Grid: endy=" << endy << "\n";
    std::cout << "Grid: cfl=" << m_cfl << "\n";
}
