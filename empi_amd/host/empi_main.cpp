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
//   --epochs-per-batch=N   decompose N epochs per device batch (default: all)
// By default every epoch of the file is decomposed together in one
// device-resident batch and the progress lines are printed afterwards.  With
// --epochs-per-batch the file is streamed: a reader thread reads the next N
// epochs while the device decomposes the current ones, the batch plan is
// built once and reused, and each batch's lines and book records are
// written, in epoch order, as soon as it is done.
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
#include <exception>
#include <string>
#include <thread>
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

// Progress lines and book records of one batch, epochs numbered from first + 1.
void write_batch(const Run& r, BookWriter& writer, const std::vector<Signal>& epochs,
                 const std::vector<BatchDecomposition::EpochResult>& results, int first) {
    for (size_t e = 0; e < epochs.size(); ++e) {
        const BatchDecomposition::EpochResult& res = results[e];
        const int epoch = first + static_cast<int>(e) + 1;
        start_line(epoch, r);
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
        writer.write(epoch, epochs[e], res.atoms);
    }
}

// Epochs in device-resident batches of per_batch (0: all of them); output
// replayed in reference order.  While a batch runs on the device, a reader
// thread reads the next one and copies it to the device (empi_plan_stage)
// when it has the current batch's shape.  A read error ends the stream after
// the epochs read before it are decomposed and written, as in the
// reference's epoch-by-epoch loop.
void run_batched(Run& r, int per_batch) {
    BookWriter writer(r.bookPath);
    int samples = -1;
    auto read_batch = [&](std::vector<Signal>& out, std::exception_ptr& error) {
        out.clear();
        try {
            while (per_batch <= 0 || static_cast<int>(out.size()) < per_batch) {
                Signal s = r.reader->read();
                if (s.samples == 0) break;
                // one workspace sized from the first epoch serves every epoch (src/empi.cpp:154-158)
                if (samples >= 0 && s.samples != samples) throw Error("internalLogicError");
                samples = s.samples;
                out.push_back(std::move(s));
            }
        } catch (...) {
            error = std::current_exception();
        }
    };
    BatchDecomposition batch;
    std::vector<Signal> epochs, next;
    std::exception_ptr error;
    read_batch(epochs, error);
    int first = 0;           // epochs written so far
    bool staged = false;     // epochs' samples are on the device already
    while (!epochs.empty()) {
        std::exception_ptr next_error;
        bool next_staged = false;
        std::thread reader;
        if (!staged) batch.prepare(epochs, r.variant, r.settings, r.dict);
        if (!error) reader = std::thread([&] {
            read_batch(next, next_error);
            if (!next_error) {
                try {
                    next_staged = batch.stage(next, r.variant, r.settings, r.dict);
                } catch (...) {
                    next_error = std::current_exception();
                }
            }
        });
        std::vector<BatchDecomposition::EpochResult> results;
        try {
            results = batch.run(epochs, r.variant, r.settings, r.dict, staged);
        } catch (...) {
            if (reader.joinable()) reader.join();
            throw;
        }
        if (reader.joinable()) reader.join();
        write_batch(r, writer, epochs, results, first);
        first += static_cast<int>(epochs.size());
        if (error) break;
        epochs.swap(next);
        error = next_error;
        staged = next_staged;
    }
    if (error) std::rethrow_exception(error);
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
    int per_batch = 0;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--per-call") == 0) per_call = true;
        if (std::strncmp(argv[i], "--epochs-per-batch=", 19) == 0) {
            per_batch = std::atoi(argv[i] + 19);
            if (per_batch <= 0) {
                std::fprintf(stderr, "%s: --epochs-per-batch needs a positive count\n", argv[0]);
                return 1;
            }
        }
        if (*argv[i] != '-' && *argv[i] != 0) config = argv[i];
    }
    if (!config) {
        std::fprintf(stderr, "USAGE: %s [--per-call] [--epochs-per-batch=N] path_to_config_file\n", argv[0]);
        return 1;
    }
    try {
        Run r = configure(config);
        if (per_call) run_per_call(r);
        else run_batched(r, per_batch);
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
