// Cartesian grid with a 1-cell halo (reference: src/Grid.h, src/Grid.cpp:4-161).
#ifndef RT2D_GRID_H
#define RT2D_GRID_H
#include "DataTypes.hh"

class Grid {
  private:
    double m_cfl = 1;

  public:
    size_t nx, ny, nxy;
    double dx, dy, dt;
    double startx, starty;
    double endx, endy;

    Grid() = delete;
    Grid(size_t nx_, size_t ny_, Coord start_, Coord end_, size_t halo = 1);
    Grid(const Grid &grid) = default;

    void SetCFL(double cfl);
    double GetCFL() const { return m_cfl; }

    size_t Index(size_t i, size_t j) const { return i + j * nx; }
    double x(size_t i) const { return startx + i * dx; }
    double y(size_t j) const { return starty + j * dy; }
    double x(double i) const { return startx + i * dx; }
    double y(double j) const { return starty + j * dy; }
    Coord xy(size_t i, size_t j) const { return Coord(x(i), y(j)); }
    Coord xy(size_t ij) const { return xy(i(ij), j(ij)); }
    double i(double x) const { return (x - startx) / dx; }
    double j(double y) const { return (y - starty) / dy; }
    size_t i(size_t ij) const { return ij % nx; }
    size_t j(size_t ij) const { return ij / nx; }
    Coord ij(const Coord &xy) const { return Coord(i(xy[1]), j(xy[2])); }
    // Frame output (reference src/Grid.cpp:109-136).
    void WriteFrametoCsv(float time, const DoubleBuffer &r, const DoubleBuffer &g, const DoubleBuffer &b,
                         const DoubleBuffer &a, std::string directory, std::string name = "");
    bool OutsideDomain(const Coord &xy) const
    {
        return xy[1] > endx || xy[1] < startx || xy[2] > endy || xy[2] < starty;
    }
    bool OutsideDomain(double i, double j) const { return i < 0 || j < 0 || i > nx - 1 || j > ny - 1; }
    // Grid::Print (src/Grid.cpp:140-153): the grid's sizes, one per line.
    void Print() const;
};
#endif
