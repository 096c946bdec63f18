// ncorr.cpp -- the drop-in ncorr class (include/cudaNcorr.hpp) over the C-ABI
// of libncorr_hip.so.  Host C++ only; no HIP header is included.
//
// Reference: source/cudaNcorr.cu.  Constructor geometry and the output bin
// list follow cudaNcorr.cu:118-157 (threeD) and 159-216 (rppi) (bin mids
// cudaNcorr.cu:26-35, output filter mid_k <= mid_i + mid_j at cudaNcorr.cu:148);
// calculateD1D2D3 replaces cudaNcorr.cu:220-430 (gridding, per-cell uploads,
// kernel launch, copy-back) with one nc_count_host call; the compactions follow
// cudaNcorr.cu:72-86 and 88-105.
#include <cmath>
#include <stdexcept>
#include <string>

#include "cudaNcorr.hpp"
#include "ncorr_hip.h"

namespace {

// Error convention of the reference: std::runtime_error carrying the failing
// call (gpuerrchk.h:9-14); bad arguments raise std::invalid_argument.
void check(nc_context *ctx, int rc, const char *what) {
    if (rc == NC_OK) return;
    std::string msg = std::string(what) + ": " + (ctx ? nc_last_error(ctx) : "no device context");
    if (rc == NC_ERR_BINS || rc == NC_ERR_ARG || rc == NC_ERR_DOMAIN) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void validate_edges(const std::vector<double> &b) {
    if (b.size() < 2) throw std::invalid_argument("ncorr: need at least two bin edges");
    if (!(b[0] >= 0.0)) throw std::invalid_argument("ncorr: first bin edge must be >= 0");
    for (size_t i = 1; i < b.size(); ++i)
        if (!(b[i] > b[i - 1])) throw std::invalid_argument("ncorr: bin edges must be strictly increasing");
}

}  // namespace

std::vector<double> ncorr::getBinsMid(const std::vector<double> &bins) {
    std::vector<double> mid;
    for (size_t i = 1; i < bins.size(); ++i) mid.push_back((bins[i] + bins[i - 1]) / 2.0);
    return mid;
}

ncorr::ncorr(double3 box, int bin_type, std::vector<double> &bins, bool isSurvey, double3 ori)
    : binType_(static_cast<binType>(bin_type)), isSurvey_(isSurvey), bins_(bins), box_(box), ori_(ori) {
    if (binType_ != threeD) throw std::runtime_error("Invalid constructor call for bin_type.");
    validate_edges(bins_);
    binsMid_ = getBinsMid(bins_);
    rMax_ = bins_.back();
    rMin_ = bins_[0];
    numBins_ = (int)bins_.size() - 1;
    numBinsCounts_ = numBins_ * numBins_ * numBins_;
    cellSize_ = {rMax_, rMax_, rMax_};
    numCells_.x = static_cast<int>(std::ceil(box_.x / cellSize_.x));
    numCells_.y = static_cast<int>(std::ceil(box_.y / cellSize_.y));
    numCells_.z = static_cast<int>(std::ceil(box_.z / cellSize_.z));
    for (int i = 0; i < numBins_; ++i)
        for (int j = i; j < numBins_; ++j)
            for (int k = j; k < numBins_; ++k)
                if (binsMid_[k] <= binsMid_[i] + binsMid_[j]) {
                    triCountsOut_.push_back(0.0);
                    triBinsOut_.push_back({binsMid_[i], binsMid_[j], binsMid_[k]});
                }
    triCountsAll_.assign(numBinsCounts_, 0.0);
}

ncorr::ncorr(double3 box, int bin_type, std::vector<double> &bins, int numMinorBins, bool isSurvey, double3 ori)
    : binType_(static_cast<binType>(bin_type)), isSurvey_(isSurvey), bins_(bins), box_(box), ori_(ori) {
    numMinorBins_ = numMinorBins;
    if (binType_ != rppi) throw std::runtime_error("Invalid constructor call for bin_type.");
    validate_edges(bins_);
    if (numMinorBins < 1) throw std::invalid_argument("ncorr: need at least one pi bin");
    numBins_ = (int)bins_.size() - 1;
    binsMid_ = getBinsMid(bins_);
    rpMax_ = bins_.back();
    rpMin_ = bins_[0];
    // unit-width pi bin edges 0, 1, ..., numMinorBins (cudaNcorr.cu:173-175)
    piBins_.resize((size_t)numMinorBins + 1);
    for (int i = 0; i <= numMinorBins; ++i) piBins_[(size_t)i] = (double)i;
    numPiBins_ = numMinorBins;
    piMax_ = piBins_.back();
    const double sqrRMax = rpMax_ * rpMax_ + piMax_ * piMax_;
    rMax_ = sqrtf((float)sqrRMax);  // FP32, as cudaNcorr.cu:187
    rMin_ = rpMin_;
    numBinsCounts_ = numBins_ * numBins_ * numBins_ * numPiBins_ * numPiBins_;
    cellSize_ = isSurvey_ ? double3{rMax_, rMax_, rMax_} : double3{rpMax_, rpMax_, piMax_};
    numCells_.x = static_cast<int>(std::ceil(box_.x / cellSize_.x));
    numCells_.y = static_cast<int>(std::ceil(box_.y / cellSize_.y));
    numCells_.z = static_cast<int>(std::ceil(box_.z / cellSize_.z));
    for (int i = 0; i < numBins_; ++i)
        for (int j = i; j < numBins_; ++j)
            for (int k = j; k < numBins_; ++k)
                if (binsMid_[k] <= binsMid_[i] + binsMid_[j])
                    for (int l = 0; l < numPiBins_; ++l)
                        for (int m = 0; m < numPiBins_; ++m) {
                            triCountsOut_.push_back(0.0);
                            triBinsOut_.push_back({binsMid_[i], binsMid_[j], binsMid_[k]});
                            piBinsOut_.push_back({piBins_[l], piBins_[m]});
                        }
    triCountsAll_.assign(numBinsCounts_, 0.0);
}

void ncorr::setDevice(int device) {
    if (device != device_) ctx_.reset();
    device_ = device;
}

nc_context *ncorr::context() {
    if (!ctx_) {
        nc_context *c = nullptr;
        int rc = nc_create(&c, device_);
        if (rc != NC_OK) throw std::runtime_error("ncorr: cannot create a HIP context on device " + std::to_string(device_));
        ctx_ = std::shared_ptr<nc_context>(c, [](nc_context *p) { nc_destroy(p); });
    }
    return ctx_.get();
}

void ncorr::calculateD1D2D3(std::vector<double4> &p1, std::vector<double4> &p2, std::vector<double4> &p3) {
    nc_context *ctx = context();
    nc_params P{};
    P.box[0] = box_.x;
    P.box[1] = box_.y;
    P.box[2] = box_.z;
    P.survey = isSurvey_ ? 1 : 0;
    P.nedges = (int)bins_.size();
    P.edges = bins_.data();
    if (binType_ == rppi) {
        P.bin_type = NC_BIN_RPPI;
        P.num_pi_bins = numPiBins_;
        P.ori[0] = ori_.x;
        P.ori[1] = ori_.y;
        P.ori[2] = ori_.z;
    }
    triCountsAll_.assign(numBinsCounts_, 0.0);
    const double *d1 = reinterpret_cast<const double *>(p1.data());
    const double *d2 = &p2 == &p1 ? d1 : reinterpret_cast<const double *>(p2.data());
    const double *d3 = &p3 == &p2 ? d2 : (&p3 == &p1 ? d1 : reinterpret_cast<const double *>(p3.data()));
    check(ctx,
          nc_count_host(ctx, &P, d1, (int64_t)p1.size(), d2, (int64_t)p2.size(), d3, (int64_t)p3.size(),
                        precision_ == mixedFP32 ? NC_PATH_MIXED : NC_PATH_PLAIN, triCountsAll_.data(), nullptr),
          "calculateD1D2D3");
    nc_timing t;
    if (nc_last_timing(ctx, &t) == NC_OK) lastMs_ = t.ms_total;
    if (binType_ == rppi)
        prepTriCountsOutRppi();
    else
        prepTriCountsOutThreeD();
}

// cudaNcorr.cu:88-105: the threeD selection, each entry followed by its
// npi^2 pi bins.
void ncorr::prepTriCountsOutRppi() {
    int iout = 0;
    const int np = numPiBins_, nb = numBins_;
    for (int i = 0; i < nb; ++i)
        for (int j = i; j < nb; ++j)
            for (int k = j; k < nb; ++k)
                if (binsMid_[k] <= binsMid_[i] + binsMid_[j])
                    for (int l = 0; l < np; ++l)
                        for (int m = 0; m < np; ++m)
                            triCountsOut_[iout++] =
                                triCountsAll_[m + l * np + k * np * np + j * np * np * nb + i * np * np * nb * nb];
}

void ncorr::prepTriCountsOutThreeD() {
    int iout = 0;
    for (int i = 0; i < numBins_; ++i)
        for (int j = i; j < numBins_; ++j)
            for (int k = j; k < numBins_; ++k)
                if (binsMid_[k] <= binsMid_[i] + binsMid_[j])
                    triCountsOut_[iout++] = triCountsAll_[k + j * numBins_ + i * numBins_ * numBins_];
}

std::vector<double3> ncorr::getTriBinsOut() const { return triBinsOut_; }

std::vector<double2> ncorr::getPiBinsOut() const { return piBinsOut_; }

std::vector<double> ncorr::getTriCountsOut() const { return triCountsOut_; }
