// Simulation configuration (reference: src/Config.hh:1-60).  The reference's
// fields keep their order so designated initialisers compile unchanged; the
// fields after initialDataType are additions of this library.
#ifndef RT2D_CONFIG_HH
#define RT2D_CONFIG_HH
#include "Utility.hh"

enum StreamingType { FlatFixed, FlatAdaptive, CurvedFixed, CurvedAdaptive, GeodesicFixed };
inline std::string StreamingName(int n)
{
    switch (n) {
    case 0: return "FlatFixed";
    case 1: return "FlatAdaptive";
    case 2: return "CurvedFixed";
    case 3: return "CurvedAdaptive";
    case 4: return "GeodesicFixed";
    default: ExitOnError("Invalid StreamingType");
    }
    return "unknown";
}
enum InitialDataType { Moments, Intensities };
enum Precision { FP64 = 0, FP32Mixed = 1 };

struct Config {
    std::string name;
    double t0 = 0;
    double simTime = 1;
    double writePeriod = 1;
    bool updateFourierHarmonics = false;
    bool keepSourceNodesActive = false;
    bool writeData = true;
    bool printSetup = true;
    bool printProgress = true;
    bool printResults = true;
    bool saveInitialData = true;
    StreamingType streamingType = StreamingType::FlatFixed;
    InitialDataType initialDataType = InitialDataType::Moments;
    // Additions: arithmetic precision of the GPU step, excision radius
    // (negative: the metric's own InsideBH, r <= 2.1 m), RK45 tolerance of the
    // geodesic tables (the reference hard-codes 1e-4), GPU ordinal.
    Precision precision = Precision::FP64;
    double excisionRadius = -1;
    double rk45Tolerance = 1e-4;
    int device = 0;
};
#endif
