// Utilities of the reference API subset used by the stream-collide path
// (reference: src/Utility.hh).  Hastings trig comes from rt_math.h.
#ifndef RT2D_UTILITY_HH
#define RT2D_UTILITY_HH
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>
#include "../csrc/rt_math.h"

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

inline double MyAtan2(double y, double x) { return rtm::my_atan2(y, x); }
inline double MySin(double x) { return rtm::my_sin(x); }
inline double MyCos(double x) { return rtm::my_cos(x); }
inline double MyAtan(double z) { return rtm::hastings_atan(z); }
#endif
