// empi-hip: command-line drop-in for empi (src/empi.cpp:49-195) on MI355X.
//
//   empi-hip [options] config_file
//
// Same configuration keys, validation and error codes, output book name
// (<dir>/<base>_<first 3 letters of MP>.b) and stdout protocol
// (START / CHANNEL / ATOM / " END" / ERROR) as empi.  Arguments starting
// with '-' are ignored for MP5 compatibility, except:
//   --per-call   run the reference loop on the host over the per-call
//                Workspace ABI (one device round trip per atom)
// By default every epoch of the file is decomposed together in one
// device-resident batch and the progress lines are printed afterwards.
#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "book.hpp"
#include "config.hpp"
#include "signal_reader.hpp"
#include "workspace.hpp"

using namespace empi::host;

namespace {

struct Run {
    empi_dict_params dict{};
    DecompositionSettings settings;
    std::unique_ptr<SignalReader> reader;
    std::string mp;           // lower-case variant name
    int variant = EMPI_SMP;
    int channelCount = 0;     // selected channels
    int builderChannels = 0;  // channels of the workspace (1 for SMP and MMP2)
    std::string bookPath;
};

// src/empi.cpp:55-141
Run configure(const char* path) {
    Configuration cfg;
    cfg.parse(path);
    Run r;
    const double energyError = std::atof(cfg.at("energyError").c_str());
    if (energyError <= 0.0 || energyError >= 1.0) throw Error("invalidEnergyErrorValue");
    r.dict.energyError = energyError;
    r.dict.scaleMin = cfg.has("minAtomScale") ? std::atof(cfg.at("minAtomScale").c_str()) : 0.0;
    r.dict.scaleMax = cfg.has("maxAtomScale") ? std::atof(cfg.at("maxAtomScale").c_str()) : INFINITY;
    r.dict.freqMax = cfg.has("maxAtomFrequency") ? std::atof(cfg.at("maxAtomFrequency").c_str()) : INFINITY;

    r.settings.iterationMax = std::atoi(cfg.at("maximalNumberOfIterations").c_str());
    if (r.settings.iterationMax <= 0) throw Error("invalidMaximalNumberOfIterations");
    r.settings.energyPercent = std::atof(cfg.at("energyPercent").c_str());
    r.settings.residualEnergy = 1 - 0.01 * r.settings.energyPercent;
    if (r.settings.residualEnergy < 0 || r.settings.residualEnergy >= 1.0) throw Error("invalidEnergyPercentValue");

    const std::string& dataFile = cfg.at("nameOfDataFile");
    if (cfg.has("numberOfSamplesInEpoch")) {
        const int epochSize = std::atoi(cfg.at("numberOfSamplesInEpoch").c_str());
        if (epochSize <= 0) throw Error("invalidNumberOfSamplesInEpoch");
        if (cfg.has("selectedEpochs")) {
            std::vector<int> epochs = parse_integer_subset(cfg.at("selectedEpochs"), INT_MAX / epochSize);
            if (epochs.empty()) throw Error("noSelectedEpochs");
            r.reader.reset(new SignalReaderForSelectedEpochs(dataFile, epochSize, epochs));
        } else {
            r.reader.reset(new SignalReaderForAllEpochs(dataFile, epochSize));
        }
    } else {
        if (cfg.has("selectedEpochs")) cfg.at("numberOfSamplesInEpoch");   // throws the missing entry
        r.reader.reset(new SignalReaderForWholeSignal(dataFile));
    }
    r.reader->fs = std::atof(cfg.at("samplingFrequency").c_str());
    if (r.reader->fs <= 0.0) throw Error("invalidSamplingFrequency");
    r.reader->channelCount = std::atoi(cfg.at("numberOfChannels").c_str());
    if (r.reader->channelCount <= 0) throw Error("invalidNumberOfChannels");
    r.reader->selectedChannels = parse_integer_subset(cfg.at("selectedChannels"), r.reader->channelCount);

    r.channelCount = static_cast<int>(r.reader->selectedChannels.size());
    r.builderChannels = r.channelCount;
    std::string mp = cfg.at("MP");
    std::transform(mp.begin(), mp.end(), mp.begin(), [](unsigned char c) { return std::tolower(c); });
    if (r.reader->selectedChannels.size() == 1 || mp == "smp") {
        r.variant = EMPI_SMP;
        r.builderChannels = 1;
    } else if (mp == "mmp1") {
        r.variant = EMPI_MMP1;
    } else if (mp == "mmp2") {
        r.variant = EMPI_MMP2;
        r.builderChannels = 1;
    } else if (mp == "mmp3") {
        r.variant = EMPI_MMP3;
    } else {
        throw Error("unsupportedDecompositionType");
    }
    std::string prefix = r.reader->path;
    const size_t slash = prefix.find_last_of('/');
    if (slash != std::string::npos) prefix = prefix.substr(slash + 1);
    prefix = prefix.substr(0, prefix.find_last_of('.'));
    mp.resize(3);
    r.mp = mp;
    r.bookPath = cfg.at("nameOfOutputDirectory") + "/" + prefix + "_" + mp + ".b";
    return r;
}

std::unique_ptr<Decomposition> make_decomposition(int variant, std::ostream& out) {
    switch (variant) {
        case EMPI_SMP: return std::unique_ptr<Decomposition>(new SmpDecomposition(out));
        case EMPI_MMP1: return std::unique_ptr<Decomposition>(new Mmp1Decomposition(out));
        case EMPI_MMP2: return std::unique_ptr<Decomposition>(new Mmp2Decomposition(out));
        default: return std::unique_ptr<Decomposition>(new Mmp3Decomposition(out));
    }
}

void start_line(int epoch, const Run& r) {
    std::cout << "START" << '\t' << epoch << '\t' << r.channelCount << '\t' << r.settings.iterationMax << '\t'
              << 100 * (1 - r.settings.residualEnergy) << std::endl;
}

// src/empi.cpp:145-164, the reference loop over the per-call Workspace ABI
void run_per_call(Run& r) {
    BookWriter writer(r.bookPath);
    WorkspaceBuilder builder(r.dict);
    std::unique_ptr<Decomposition> decomposition = make_decomposition(r.variant, std::cout);
    std::unique_ptr<Workspace> ws;
    int epoch = 0;
    while (true) {
        Signal signal = r.reader->read();
        if (signal.samples == 0) break;
        if (!ws)
            ws.reset(builder.prepareWorkspace(r.reader->fs, r.builderChannels, signal.samples,
                                              decomposition->constraint()));
        start_line(++epoch, r);
        ChannelAtoms result = decomposition->compute(r.settings, ws.get(), signal);
        writer.write(epoch, signal, result);
    }
    writer.close();
}

// every epoch in one device-resident batch; output replayed in reference order
void run_batched(Run& r) {
    BookWriter writer(r.bookPath);
    std::vector<Signal> epochs;
    while (true) {
        Signal s = r.reader->read();
        if (s.samples == 0) break;
        // one workspace sized from the first epoch serves every epoch (src/empi.cpp:154-158)
        if (!epochs.empty() && s.samples != epochs.front().samples) throw Error("internalLogicError");
        epochs.push_back(std::move(s));
    }
    BatchDecomposition batch;
    std::vector<BatchDecomposition::EpochResult> results =
        batch.run(epochs, r.variant, r.settings, r.dict);
    for (size_t e = 0; e < epochs.size(); ++e) {
        const BatchDecomposition::EpochResult& res = results[e];
        start_line(static_cast<int>(e) + 1, r);
        const bool failed = res.status == EMPI_ENERGY_MAP_IS_EMPTY;
        if (r.variant == EMPI_SMP) {
            for (int c = 0; c < epochs[e].channels; ++c) {
                std::cout << "CHANNEL\t" << c << std::endl;
                const std::vector<double>& en = res.energies[c];
                if (std::isnan(en[0]) || en[0] == 0) throw Error("signalIsEmpty");
                const bool chan_failed = failed && res.counts[c] < r.settings.iterationMax &&
                                         !(en[res.counts[c]] / en[0] <= r.settings.residualEnergy);
                BatchDecomposition::print_progress(std::cout, en, res.counts[c], res.atomCount,
                                                   r.settings.iterationMax, chan_failed);
                if (chan_failed) throw Error("energyMapIsEmpty");
            }
        } else {
            const std::vector<double>& en = res.energies[0];
            if (std::isnan(en[0]) || en[0] == 0) throw Error("signalIsEmpty");
            BatchDecomposition::print_progress(std::cout, en, res.counts[0], res.atomCount, r.settings.iterationMax,
                                               failed);
            if (failed) throw Error("energyMapIsEmpty");
        }
        if (res.status != EMPI_OK) throw Error(empi_strerror(res.status));
        writer.write(static_cast<int>(e) + 1, epochs[e], res.atoms);
    }
    writer.close();
}

void fail(const char* message) {
    std::cout << "ERROR\t" << message << std::endl;
    std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv) {
    const char* config = nullptr;
    bool per_call = false;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--per-call") == 0) per_call = true;
        if (*argv[i] != '-' && *argv[i] != 0) config = argv[i];
    }
    if (!config) {
        std::fprintf(stderr, "USAGE: %s [--per-call] path_to_config_file\n", argv[0]);
        return 1;
    }
    try {
        Run r = configure(config);
        if (per_call) run_per_call(r);
        else run_batched(r);
        std::puts(" END");
    } catch (const Error& e) {
        fail(e.what());
    } catch (const std::bad_alloc&) {
        fail("insufficientMemory");
    } catch (...) {
        fail("internalError");
    }
    return EXIT_SUCCESS;
}
