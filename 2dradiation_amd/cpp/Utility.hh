// Utilities of the reference API subset used by the stream-collide path
// (reference: src/Utility.hh).  Hastings trig comes from rt_math.h.
#ifndef RT2D_UTILITY_HH
#define RT2D_UTILITY_HH
#include <cerrno>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>
#include "../csrc/rt_math.h"

// PARALLEL_FOR(N) (src/ControlFlow.hh:23-24) for user loops: an OpenMP
// collapse(N) loop when the user program is built with -fopenmp.
#define RT2D_PRAGMA(x) _Pragma(#x)
#define PARALLEL_FOR(N) RT2D_PRAGMA(omp parallel for collapse(N) schedule(dynamic, 50))

// ExitOnError (src/Utility.hh:16-20): print and exit with errno.
inline void ExitOnError(std::string msg = "")
{
    std::cout << "ERROR: " + msg + "\n";
    exit(errno ? errno : 1);
}

// 64-byte aligned vectors, as the reference's DoubleBuffer (src/Utility.hh:33-76).
template <class T>
struct AlignedArrayAllocator {
    using value_type = T;
    AlignedArrayAllocator() = default;
    template <class U>
    AlignedArrayAllocator(const AlignedArrayAllocator<U> &) {}
    T *allocate(size_t n)
    {
        size_t bytes = ((n * sizeof(T) + 63) / 64) * 64;
        return static_cast<T *>(std::aligned_alloc(64, bytes ? bytes : 64));
    }
    void deallocate(T *p, size_t) { std::free(p); }
    bool operator==(const AlignedArrayAllocator &) const { return true; }
    bool operator!=(const AlignedArrayAllocator &) const { return false; }
};
using IntBuffer = std::vector<int, AlignedArrayAllocator<int>>;
using DoubleBuffer = std::vector<double, AlignedArrayAllocator<double>>;
using FloatBuffer = std::vector<float, AlignedArrayAllocator<float>>;

// IntegerPow (src/Utility.hh:96-105): a*(a*(...*(a*1))).
template <int N>
inline double IntegerPow(double a) { return a * IntegerPow<N - 1>(a); }
template <>
inline double IntegerPow<0>(double) { return 1; }

// Output root of the frame files and Log.txt (src/ControlFlow.hh:18), relative
// to the working directory, as in the reference.
#define OUTPUTDIR (std::string) "../output/"

// Fixed-point number text of the reference's Format (src/Utility.hh:511-540):
// a leading space for n >= 0 (and for +0), none for negatives and -0; one
// more space per missing digit below maxDigits (up to 3); `precision`
// decimals.
inline std::string Format(const double n, const int precision = 6, const bool noSpace = false, const int maxDigits = 1)
{
    std::string out;
    if (!noSpace && !std::signbit(n)) out = " ";
    if (maxDigits > 2 && std::abs(n) < 100) out += " ";
    if (maxDigits > 1 && std::abs(n) < 10) out += " ";
    std::ostringstream ss;
    ss.precision(precision);
    ss << std::fixed << n;
    return out + ss.str();
}

// Creates `path` (relative to the working directory) with its parents
// (src/Utility.hh:594-605); false if it existed or path is empty.
inline bool CreateDirectory(const std::string &path)
{
    if (path.empty()) return false;
    return std::filesystem::create_directories(std::filesystem::current_path() / path);
}

inline double MyAtan2(double y, double x) { return rtm::my_atan2(y, x); }
// Arc cosine of the reference (MyAcos, src/Utility.hh:453-471 with
// APPROXIMATION_ORDER 9): the Hastings (1955) 4th-order arcsine
// asin(x) = pi/2 - sqrt(1-x) (a0 + a1 x + ... + a4 x^4) for x in [0, 1], odd
// in x; the argument and the result pass through float, as there.
inline double MyAsinHastings4(double x)
{
    const double s = (x > 0) - (x < 0);
    x = std::abs(x);
    const double p = 1.57078786 + x * (-0.21412453 + x * (0.08466649 + x * (-0.03575663 + x * 0.00864884)));
    return s * (rtm::kPi / 2 - std::sqrt(1.0 - x) * p);
}
inline float MyAcos(float x)
{
    return x >= 0 ? (float)(rtm::kPi / 2 - MyAsinHastings4(x)) : (float)(rtm::kPi / 2 + MyAsinHastings4(-x));
}
inline double MyAcos(double x) { return MyAcos((float)x); }
inline double MySin(double x) { return rtm::my_sin(x); }
inline double MyCos(double x) { return rtm::my_cos(x); }
inline double MyAtan(double z) { return rtm::hastings_atan(z); }
#endif
