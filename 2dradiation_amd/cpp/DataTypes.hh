// Tensor types of the reference API subset used on the stream-collide path
// (reference: src/DataTypes.hh:37-213, 620-635).  Components are indexed like
// the reference: Coord/Tensor2 use [1],[2]; Tensor3/Tensor3x3 use [0..2].
#ifndef RT2D_DATATYPES_HH
#define RT2D_DATATYPES_HH
#include "Utility.hh"

struct rank2Indices { int i, j; };

// Two-component spatial vectors (Coord: positions, Tensor2: directions and
// velocities) with the reference's 1-based components and componentwise
// arithmetic (src/DataTypes.hh:37-213).
template <class V>
struct Vec2Base {
    double data[2];
    double &operator[](int k) { return data[k - 1]; }
    const double &operator[](int k) const { return data[k - 1]; }
    double EuklNormSquared() const { return data[0] * data[0] + data[1] * data[1]; }
    double EuklNorm() const { return std::sqrt(EuklNormSquared()); }
    double Phi() const { return rtm::my_phi(data[0], data[1]); }
    V EuklNormalized() const
    {
        const double n = EuklNorm();
        return V(data[0] / n, data[1] / n);
    }
    V operator+(const V &o) const { return V(data[0] + o.data[0], data[1] + o.data[1]); }
    V operator-(const V &o) const { return V(data[0] - o.data[0], data[1] - o.data[1]); }
    V operator-() const { return V(-data[0], -data[1]); }
    V operator*(double s) const { return V(data[0] * s, data[1] * s); }
    V operator/(double s) const { return V(data[0] / s, data[1] / s); }
    friend V operator*(double s, const V &v) { return V(s * v.data[0], s * v.data[1]); }
    static double Dot(const V &a, const V &b) { return a[1] * b[1] + a[2] * b[2]; }
    static double Cross(const V &a, const V &b) { return a[1] * b[2] - b[1] * a[2]; }
    friend std::ostream &operator<<(std::ostream &os, const V &v) { return os << "(" << v[1] << "," << v[2] << ")"; }
};
struct Coord : Vec2Base<Coord> {
    Coord(double v = 0) { data[0] = data[1] = v; }
    Coord(double x, double y) { data[0] = x; data[1] = y; }
};
struct Tensor2 : Vec2Base<Tensor2> {
    Tensor2(double v = 0) { data[0] = data[1] = v; }
    Tensor2(double x, double y) { data[0] = x; data[1] = y; }
};

struct Tensor3 {
    double data[3];
    Tensor3(double v = 0) { data[0] = data[1] = data[2] = v; }
    Tensor3(double t, double x, double y) { data[0] = t; data[1] = x; data[2] = y; }
    double &operator[](int k) { return data[k]; }
    const double &operator[](int k) const { return data[k]; }
    Tensor3 operator+(const Tensor3 &o) const { return Tensor3(data[0] + o[0], data[1] + o[1], data[2] + o[2]); }
    Tensor3 operator-(const Tensor3 &o) const { return Tensor3(data[0] - o[0], data[1] - o[1], data[2] - o[2]); }
    Tensor3 operator*(double s) const { return Tensor3(data[0] * s, data[1] * s, data[2] * s); }
    Tensor3 operator/(double s) const { return Tensor3(data[0] / s, data[1] / s, data[2] / s); }
    friend Tensor3 operator*(double s, const Tensor3 &v) { return Tensor3(s * v[0], s * v[1], s * v[2]); }
    friend std::ostream &operator<<(std::ostream &os, const Tensor3 &v)
    {
        return os << "(" << v[0] << "," << v[1] << "," << v[2] << ")";
    }
};

struct Tensor3x3 {
    double data[9];
    Tensor3x3(double v = 0) { for (double &x : data) x = v; }
    Tensor3x3(double tt, double tx, double ty, double xt, double xx, double xy, double yt, double yx,
              double yy)
        : data{tt, tx, ty, xt, xx, xy, yt, yx, yy} {}
    double &operator[](const rank2Indices &k) { return data[k.i * 3 + k.j]; }
    const double &operator[](const rank2Indices &k) const { return data[k.i * 3 + k.j]; }
    // 3x3 inverse by the adjugate with cyclic cofactors; the reference calls
    // Eigen's fixed-size inverse here (src/DataTypes.hh:417-425).
    Tensor3x3 Invert() const
    {
        auto cof = [this](int i, int j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return data[3 * i1 + j1] * data[3 * i2 + j2] - data[3 * i1 + j2] * data[3 * i2 + j1];
        };
        double c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
        double det = (c0 * data[0] + c1 * data[3]) + c2 * data[6];
        double id = 1.0 / det;
        Tensor3x3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r.data[3 * i + j] = cof(j, i) * id;
        return r;
    }
};

struct rank3Indices { int i, j, k; };

// Spatial 2x2 and 2x2x2 tensors with the reference's 1-based indices
// ({1,1}..{2,2}), and the 3x3x3 spacetime tensor with 0-based ones, as the
// metric getters return them (src/DataTypes.hh:296-360, 450-600).
struct Tensor2x2 {
    double data[4];
    Tensor2x2(double v = 0) { data[0] = data[1] = data[2] = data[3] = v; }
    Tensor2x2(double xx, double xy, double yx, double yy) : data{xx, xy, yx, yy} {}
    double &operator[](const rank2Indices &k) { return data[(k.i - 1) * 2 + (k.j - 1)]; }
    const double &operator[](const rank2Indices &k) const { return data[(k.i - 1) * 2 + (k.j - 1)]; }
    Tensor2x2 operator*(const Tensor2x2 &o) const
    {
        return Tensor2x2(data[0] * o.data[0] + data[1] * o.data[2], data[0] * o.data[1] + data[1] * o.data[3],
                         data[2] * o.data[0] + data[3] * o.data[2], data[2] * o.data[1] + data[3] * o.data[3]);
    }
    Tensor2 operator*(const Tensor2 &v) const
    {
        return Tensor2(data[0] * v[1] + data[1] * v[2], data[2] * v[1] + data[3] * v[2]);
    }
    Tensor2x2 Invert() const
    {
        const double id = 1.0 / (data[0] * data[3] - data[1] * data[2]);
        return Tensor2x2(data[3] * id, -data[1] * id, -data[2] * id, data[0] * id);
    }
    friend std::ostream &operator<<(std::ostream &os, const Tensor2x2 &v)
    {
        return os << "(" << v.data[0] << "," << v.data[1] << "|" << v.data[2] << "," << v.data[3] << ")";
    }
};
struct Tensor2x2x2 {
    double data[8];
    Tensor2x2x2(double v = 0) { for (double &x : data) x = v; }
    Tensor2x2x2(double xxx, double xxy, double xyx, double xyy, double yxx, double yxy, double yyx, double yyy)
        : data{xxx, xxy, xyx, xyy, yxx, yxy, yyx, yyy} {}
    double &operator[](const rank3Indices &k) { return data[(k.i - 1) * 4 + (k.j - 1) * 2 + (k.k - 1)]; }
    const double &operator[](const rank3Indices &k) const { return data[(k.i - 1) * 4 + (k.j - 1) * 2 + (k.k - 1)]; }
    friend std::ostream &operator<<(std::ostream &os, const Tensor2x2x2 &v)
    {
        os << "(" << v.data[0] << "," << v.data[1] << "|" << v.data[2] << "," << v.data[3] << ")";
        return os << "(" << v.data[4] << "," << v.data[5] << "|" << v.data[6] << "," << v.data[7] << ")";
    }
};
struct Tensor3x3x3 {
    double data[27];
    Tensor3x3x3(double v = 0) { for (double &x : data) x = v; }
    // components in index order [a][b][c], a slowest
    template <class... A>
    Tensor3x3x3(double first, A... rest) : data{first, (double)rest...}
    {
        static_assert(sizeof...(rest) == 26, "Tensor3x3x3 takes 27 components");
    }
    double &operator[](const rank3Indices &k) { return data[k.i * 9 + k.j * 3 + k.k]; }
    const double &operator[](const rank3Indices &k) const { return data[k.i * 9 + k.j * 3 + k.k]; }
};

// Lorentz boost of velocity u (|u| < 1) as a 3x3 (t, x, y) matrix, the
// identity for |u|^2 < 1e-10 (reference: BoostMatrix, src/SpecialMath.cpp:109-118).
inline Tensor3x3 BoostMatrix(const Tensor2 &u)
{
    const double b2 = Tensor2::Dot(u, u);
    const double g = 1.0 / std::sqrt(1.0 - b2);
    if (b2 < 1e-10) return Tensor3x3(1, 0, 0, 0, 1, 0, 0, 0, 1);
    const double k = (g - 1) / b2;
    return Tensor3x3(g, -g * u[1], -g * u[2], -g * u[1], 1 + k * u[1] * u[1], k * u[1] * u[2], -g * u[2],
                     k * u[1] * u[2], 1 + k * u[2] * u[2]);
}

// Sorted (input, output) table with clamped linear interpolation
// (LookUpTable, src/DataTypes.hh:637-700): Add keeps the inputs increasing
// and ignores an input already present; Evaluate returns the first/last
// output outside the range, the stored output at an exact input, and the
// lerp between the neighbours otherwise.
template <class T>
struct LookUpTable {
    int size = 0;
    std::vector<T> inputs;
    std::vector<T> outputs;

    void Add(T input, T output)
    {
        auto it = std::lower_bound(inputs.begin(), inputs.end(), input);
        if (it != inputs.end() && !(input < *it)) return;
        const auto k = it - inputs.begin();
        inputs.insert(it, input);
        outputs.insert(outputs.begin() + k, output);
        size = (int)inputs.size();
    }
    T Evaluate(T x) const
    {
        if (size == 0) ExitOnError("Trying to read from empty look up table.");
        if (x <= inputs[0] || size == 1) return outputs[0];
        if (x >= inputs[size - 1]) return outputs[size - 1];
        const auto hi = std::lower_bound(inputs.begin(), inputs.end(), x);
        const size_t i1 = (size_t)(hi - inputs.begin());
        if (inputs[i1] == x) return outputs[i1];
        const size_t i0 = i1 - 1;
        const T t = (x - inputs[i0]) / (inputs[i1] - inputs[i0]);
        return outputs[i0] + t * (outputs[i1] - outputs[i0]);
    }
};

// RotationMatrix (src/DataTypes.hh:620-635).
struct RotationMatrix {
    double angle;
    explicit RotationMatrix(double a) : angle(a) {}
    RotationMatrix Inverse() const { return RotationMatrix(std::fmod(-angle + 2.0 * rtm::kPi, 2.0 * rtm::kPi)); }
};
inline Tensor2 operator*(const RotationMatrix &m, const Tensor2 &t)
{
    double s = MySin(m.angle), c = MyCos(m.angle);
    return Tensor2(t[1] * c - t[2] * s, t[1] * s + t[2] * c);
}
#endif
