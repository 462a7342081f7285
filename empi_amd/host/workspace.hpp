// The reference's operator API (inc/classes.hpp:61-107, inc/mmp.hpp:11-31)
// over the C ABI of the MI355X engine:
//   Workspace / WorkspaceBuilder   per-call empi_ws_* (reference granularity)
//   Decomposition, SmpDecomposition, Mmp1/2/3Decomposition
//                                  the reference's MP loop on the host
//   BatchDecomposition             every epoch of a run in one device-
//                                  resident empi_plan (the fast path)
// Progress lines (ATOM / CHANNEL) go to the given stream exactly as the
// reference prints them (src/classes.cpp:41-48, 83).
#pragma once

#include <cstddef>
#include <memory>
#include <ostream>
#include <vector>

#include "book.hpp"
#include "empi_hip.h"

namespace empi {
namespace host {

struct DecompositionSettings {
    int iterationMax = 0;
    double residualEnergy = 0;   // 1 - 0.01 * energyPercent
    double energyPercent = 0;    // as configured (the batched plan derives residualEnergy from it)
};

// Throws Error(empi_strerror(st)) for st != EMPI_OK.
void check(empi_status st);

class Workspace {
public:
    Workspace(const empi_dict_params& dict, double fs, int channels, int samples, int constraint);
    ~Workspace();
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;

    void compute(const Signal& signal);
    std::vector<Atom> findBestMatch();
    size_t getAtomCount() const;
    void subtractAtom(const Atom& atom, double* residual, int channel);
    static void subtractAtomFromSignal(Atom& atom, double fs, double* signal, int samples, bool fit);

    int channels() const { return channels_; }

private:
    empi_ws* ws_ = nullptr;
    int channels_ = 0, samples_ = 0;
    double fs_ = 0;
};

class WorkspaceBuilder {
public:
    explicit WorkspaceBuilder(const empi_dict_params& dict) : dict_(dict) {}
    Workspace* prepareWorkspace(double fs, int channels, int samples, int constraint) const;
    const empi_dict_params& params() const { return dict_; }

private:
    empi_dict_params dict_;
};

class Decomposition {
public:
    explicit Decomposition(int constraint, std::ostream& out) : constraint_(constraint), out_(out) {}
    virtual ~Decomposition() = default;
    virtual ChannelAtoms compute(const DecompositionSettings& settings, Workspace* ws, const Signal& signal);
    int constraint() const { return constraint_; }

protected:
    int constraint_;
    std::ostream& out_;
};

class SmpDecomposition : public Decomposition {
public:
    explicit SmpDecomposition(std::ostream& out) : Decomposition(EMPI_CONSTRAINT_NONE, out) {}
    ChannelAtoms compute(const DecompositionSettings& settings, Workspace* ws, const Signal& signal) override;
};

class Mmp1Decomposition : public Decomposition {
public:
    explicit Mmp1Decomposition(std::ostream& out) : Decomposition(EMPI_CONSTRAINT_MMP1, out) {}
};

class Mmp2Decomposition : public Decomposition {
public:
    explicit Mmp2Decomposition(std::ostream& out) : Decomposition(EMPI_CONSTRAINT_NONE, out) {}
    ChannelAtoms compute(const DecompositionSettings& settings, Workspace* ws, const Signal& signal) override;
};

class Mmp3Decomposition : public Decomposition {
public:
    explicit Mmp3Decomposition(std::ostream& out) : Decomposition(EMPI_CONSTRAINT_NONE, out) {}
};

// All epochs at once on the device; emits the same progress lines per epoch
// afterwards.  Results: per epoch, per channel atoms; status per epoch.
class BatchDecomposition {
public:
    struct EpochResult {
        empi_status status = EMPI_OK;
        ChannelAtoms atoms;
        std::vector<std::vector<double>> energies;   // [row][iteration + 1]
        std::vector<int> counts;
        size_t atomCount = 0;
    };
    // Decomposes the epochs as one device batch.  The plan (dictionaries,
    // device buffers, captured graph) is kept and reused by the next call
    // with the same epoch count and shapes, so a stream of equal chunks
    // builds it once.
    // staged: the epochs' samples are already on the device (stage() returned
    // true for them, and no other run() came between).
    std::vector<EpochResult> run(const std::vector<Signal>& epochs, int variant,
                                 const DecompositionSettings& settings, const empi_dict_params& dict,
                                 bool staged = false);
    // Builds (or keeps) the plan for these epochs; run() then reuses it.
    void prepare(const std::vector<Signal>& epochs, int variant, const DecompositionSettings& settings,
                 const empi_dict_params& dict);
    // Copies the next chunk's samples to the device while run() decomposes the
    // current one (called from the reader thread, after prepare() for the
    // current chunk).  Returns false, staging nothing, when the chunk does not
    // fit the current plan; run() of those epochs then stages them itself.
    bool stage(const std::vector<Signal>& epochs, int variant, const DecompositionSettings& settings,
               const empi_dict_params& dict);
    // ATOM lines of one progress row (src/classes.cpp:41-48), the failing
    // search's line included when status is energyMapIsEmpty.
    static void print_progress(std::ostream& out, const std::vector<double>& energies, int count, size_t atomCount,
                               int iterationMax, bool failed_search);

private:
    std::unique_ptr<empi_plan, void (*)(empi_plan*)> plan_{nullptr, empi_plan_destroy};
    std::vector<empi_item> items_;   // the items plan_ was built for
};

}  // namespace host
}  // namespace empi
