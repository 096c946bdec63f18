// cudaNcorr.hpp -- drop-in replacement of the reference's public header
// (include/cudaNcorr.hpp:10-68): same class name, constructors and methods,
// same output order.  The counts come from libncorr_hip.so (HIP kernels for
// MI355X) through the C-ABI in ncorr_hip.h.
#ifndef _CUDANCORR_HPP_
#define _CUDANCORR_HPP_

#include <memory>
#include <random>
#include <string>
#include <vector>

#include "nc_vector_types.hpp"

struct nc_context;

class ncorr {
   public:
    enum binType { threeD, rppi, smu };
    // Arithmetic of the triplet kernel (extension: the reference has FP64 only).
    enum precision { plainFP64 = 0, mixedFP32 = 1 };

    // threeD constructor (cudaNcorr.cu:118-157).  Throws std::runtime_error if
    // bin_type != threeD, and std::invalid_argument for edges that are not
    // >= 0 and strictly increasing (undefined behaviour in the reference).
    ncorr(double3 box, int bin_type, std::vector<double> &bins, bool isSurvey, double3 ori = double3{0.0, 0.0, 0.0});
    // rppi constructor (cudaNcorr.cu:159-216): bins are rp edges, pi bins are
    // 0, 1, ..., numMinorBins.  Throws std::runtime_error("Invalid constructor
    // call for bin_type.") for bin_type != rppi, as the reference, and
    // std::invalid_argument for bad edges or numMinorBins < 1.
    ncorr(double3 box, int bin_type, std::vector<double> &bins, int numMinorBins, bool isSurvey,
          double3 ori = double3{0.0, 0.0, 0.0});

    // Ordered-triplet histogram of (p1, p2, p3) (cudaNcorr.cu:220-430).  The
    // inputs are only read.  Passing the same vector object twice lets the
    // engine use the (p2,p3) exchange symmetry.
    void calculateD1D2D3(std::vector<double4> &p1, std::vector<double4> &p2, std::vector<double4> &p3);

    std::vector<double3> getTriBinsOut() const;
    std::vector<double2> getPiBinsOut() const;
    std::vector<double> getTriCountsOut() const;

    // Extensions.
    void setPrecision(precision p) { precision_ = p; }
    void setDevice(int device);
    // Dense histogram of the last call (the reference's triCountsAll_: nb^3,
    // or nb^3 * npi^2 for rppi).
    std::vector<double> getTriCountsAll() const { return triCountsAll_; }
    // Device milliseconds of the last call: cell-list build + triplet kernel.
    double lastDeviceMs() const { return lastMs_; }

   private:
    binType binType_;
    bool isSurvey_;
    std::vector<double> bins_, binsMid_;
    double rMin_, rMax_;
    double rpMin_ = 0.0, rpMax_ = 0.0, piMax_ = 0.0;
    std::vector<double> piBins_;
    int numBins_, numBinsCounts_;
    int numMinorBins_ = 0, numPiBins_ = 0;
    double3 box_, cellSize_, ori_;
    int3 numCells_;
    std::vector<double3> triBinsOut_;
    std::vector<double2> piBinsOut_;
    std::vector<double> triCountsAll_;
    std::vector<double> triCountsOut_;
    precision precision_ = plainFP64;
    int device_ = 0;
    double lastMs_ = 0.0;
    std::shared_ptr<nc_context> ctx_;

    std::vector<double> getBinsMid(const std::vector<double> &bins);
    void prepTriCountsOutThreeD();
    void prepTriCountsOutRppi();
    nc_context *context();
};

#endif
