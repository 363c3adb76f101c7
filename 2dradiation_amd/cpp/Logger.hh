// Run log of a simulation (reference: src/Logger.hh:7-127).  Radiation owns
// one as its public member `logger`; RunSimulation fills it and, with
// Config::writeData, writes OUTPUTDIR/<name>/Log.txt.  The time
// measurements are the ones this library takes (the reference lists its
// profiler's scopes): "Total Time" of the step loop and the CSV frame writer
// under the reference's own scope name, so that ComputationTime() is the loop
// time without frame output in both.
#ifndef RT2D_LOGGER_HH
#define RT2D_LOGGER_HH
#include <ctime>
#include <iomanip>
#include <sstream>
#include "Metric.h"
#include "Stencil.h"

class Logger {
  public:
    // scope name of the frame writer in the reference's profiler
    static constexpr const char *kFrameWriterScope =
        "void Grid::WriteFrametoCsv(float, const DoubleBuffer&, const DoubleBuffer&, const DoubleBuffer&, "
        "const DoubleBuffer&, std::string, std::string)";

    Stencil &intensityStencil;
    Stencil &streamingStencil;
    Metric &metric;
    int timeSteps = 0;
    double simTime = 0;
    int maxItterationCount = 0;
    double averageItterationCount = 0;

    std::string name;
    std::string date;
    std::string directoryPath;
    std::vector<std::string> timeNames;
    std::vector<double> timeMeasurements;

    Logger(Stencil &intensityStencil_, Stencil &streamingStencil_, Metric &metric_)
        : intensityStencil(intensityStencil_), streamingStencil(streamingStencil_), metric(metric_)
    {
    }

    // Step count ceil(simTime/dt), the simulated time rounded to whole steps,
    // the creation date and the output directory.
    void SetValues(std::string name_, double simTime_)
    {
        name = name_;
        timeSteps = (int)std::ceil(simTime_ / metric.grid.dt);
        simTime = timeSteps * metric.grid.dt;
        const std::time_t t = std::time(nullptr);
        std::ostringstream oss;
        oss << std::put_time(std::localtime(&t), "%H.%M.%S - %d.%m.%Y");
        date = oss.str();
        directoryPath = OUTPUTDIR + name;
    }

    void AddTimeMeasurement(std::string timeName, double seconds)
    {
        timeNames.push_back(timeName);
        timeMeasurements.push_back(seconds);
    }
    double TotalTime() const { return Lookup("Total Time"); }
    double WritingTime() const { return Lookup(kFrameWriterScope); }
    double ComputationTime() const { return TotalTime() - WritingTime(); }

    void Write() const
    {
        CreateDirectory(directoryPath);
        std::ofstream file(directoryPath + "/Log.txt");
        const Grid &g = metric.grid;
        file << "Creation Date: " << date << std::endl << std::endl;
        file << "Spacetime: " << metric.Name() << std::endl;
        file << "Black Hole  Mass,   M = " << metric.m << std::endl;
        file << "Black Hole  Spin,   a = " << metric.a << std::endl << std::endl;
        file << "Grid Structure:" << std::endl;
        file << "Simulation Time = " << simTime << std::endl;
        file << "(startx,starty) = (" << g.startx << "," << g.starty << ")" << std::endl;
        file << "(endx  ,  endy) = (" << g.endx << "," << g.endy << ")" << std::endl;
        file << "(Nx    ,    Ny) = (" << g.nx << "," << g.ny << ")" << std::endl;
        file << "Nt         = " << timeSteps << std::endl;
        file << "dx         = " << g.dx << std::endl;
        file << "dy         = " << g.dy << std::endl;
        file << "dt         = " << g.dt << std::endl;
        file << "cfl        = " << g.GetCFL() << std::endl << std::endl;
        file << "Stencil Properties:" << std::endl;
        file << "Intensity Stencil: nDir          = " << intensityStencil.nDir << std::endl;
        file << "Intensity Stencil: nOrder        = " << intensityStencil.nOrder << std::endl;
        file << "Intensity Stencil: nCoefficients = " << intensityStencil.nCoefficients << std::endl;
        file << "Intensity Stencil: max Interpolation Error = " << 100 * intensityStencil.maxInterpolationError << "%"
             << std::endl;
        file << "Intensity Stencil: sigma max               = " << intensityStencil.sigmaMax << std::endl;
        file << "Intensity Stencil: relative flux max       = " << intensityStencil.relativeFluxMax << std::endl;
        file << "Streaming Stencil: nDir          = " << streamingStencil.nDir << std::endl;
        file << "Streaming Stencil: nOrder        = " << streamingStencil.nOrder << std::endl;
        file << "Streaming Stencil: nCoefficients = " << streamingStencil.nCoefficients << std::endl << std::endl;
        file << "Lambda Itteration:" << std::endl;
        file << "Max Lambda Itteration Count     = " << maxItterationCount << std::endl;
        file << "Average Lambda Itteration Count = " << averageItterationCount << std::endl << std::endl;
        file << "Time measurements:" << std::endl;
        for (size_t i = 0; i < timeMeasurements.size(); i++)
            file << timeNames[i] << ": " << timeMeasurements[i] << "s" << std::endl;
        file << "Computation Time: " << ComputationTime() << "s" << std::endl;
        file << "Benchmark: " << g.nxy * timeSteps / (1e6 * ComputationTime()) << " MLUPS" << std::endl;
    }

  private:
    double Lookup(const std::string &key) const
    {
        for (size_t i = 0; i < timeNames.size(); i++)
            if (timeNames[i] == key) return timeMeasurements[i];
        return 0;
    }
};
#endif
