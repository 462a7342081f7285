#include "workspace.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "config.hpp"

namespace empi {
namespace host {

void check(empi_status st) {
    if (st != EMPI_OK) throw Error(empi_strerror(st));
}

namespace {

Atom from_abi(const empi_atom& a) {
    Atom o;
    std::memcpy(o.params, a.params, sizeof o.params);
    o.scale = a.scale;
    o.freq = a.freq;
    o.time = a.time;
    return o;
}

empi_atom to_abi(const Atom& a) {
    empi_atom o{};
    std::memcpy(o.params, a.params, sizeof o.params);
    o.scale = a.scale;
    o.freq = a.freq;
    o.time = a.time;
    return o;
}

}  // namespace

// ------------------------------------------------------------- Workspace

Workspace::Workspace(const empi_dict_params& dict, double fs, int channels, int samples, int constraint)
    : channels_(channels), samples_(samples), fs_(fs) {
    check(empi_ws_create(&dict, fs, channels, samples, constraint, &ws_));
}

Workspace::~Workspace() { empi_ws_destroy(ws_); }

void Workspace::compute(const Signal& signal) {
    if (signal.channels != channels_ || signal.samples != samples_) throw Error("internalLogicError");
    check(empi_ws_compute(ws_, signal.data.data()));
}

std::vector<Atom> Workspace::findBestMatch() {
    std::vector<empi_atom> out(static_cast<size_t>(channels_));
    check(empi_ws_find_best(ws_, out.data()));
    std::vector<Atom> atoms;
    for (const empi_atom& a : out) atoms.push_back(from_abi(a));
    return atoms;
}

size_t Workspace::getAtomCount() const { return empi_ws_atom_count(ws_); }

void Workspace::subtractAtom(const Atom& atom, double* residual, int channel) {
    empi_atom a = to_abi(atom);
    check(empi_ws_subtract(ws_, &a, channel, residual));
}

void Workspace::subtractAtomFromSignal(Atom& atom, double fs, double* signal, int samples, bool fit) {
    empi_atom a = to_abi(atom);
    check(empi_subtract_atom_from_signal(fs, &a, signal, samples, fit ? 1 : 0));
    std::memcpy(atom.params, a.params, sizeof atom.params);
}

Workspace* WorkspaceBuilder::prepareWorkspace(double fs, int channels, int samples, int constraint) const {
    return new Workspace(dict_, fs, channels, samples, constraint);
}

// --------------------------------------------------------- Decomposition

// src/classes.cpp:26-73
ChannelAtoms Decomposition::compute(const DecompositionSettings& settings, Workspace* ws, const Signal& signal) {
    ws->compute(signal);
    const int C = signal.channels;
    Signal residue = signal;
    ChannelAtoms result(static_cast<size_t>(C));
    const size_t atomCount = ws->getAtomCount();
    const double totalEnergy = signal.energy();
    double residueEnergy = totalEnergy;
    if (totalEnergy == 0) throw Error("signalIsEmpty");
    for (int iteration = 1; iteration <= settings.iterationMax; ++iteration) {
        const double energyProgress = 100.0 * (1.0 - residueEnergy / totalEnergy);
        const double totalProgress = std::max(100.0 * (iteration - 1) / settings.iterationMax, energyProgress);
        out_ << "ATOM\t" << (iteration - 1) << '\t' << atomCount << '\t' << energyProgress << '\t' << totalProgress
             << std::endl;
        std::vector<Atom> best = ws->findBestMatch();
        for (int c = 0; c < C; ++c) result[c].push_back(best[c]);
        if (iteration == settings.iterationMax) break;
        for (int c = 0; c < C; ++c) ws->subtractAtom(best[c], residue.channel(c), c);
        residueEnergy = residue.energy();
        if (residueEnergy / totalEnergy <= settings.residualEnergy) break;
    }
    return result;
}

// src/classes.cpp:77-87: channels one after another on a 1-channel workspace
ChannelAtoms SmpDecomposition::compute(const DecompositionSettings& settings, Workspace* ws, const Signal& signal) {
    ChannelAtoms result;
    for (int c = 0; c < signal.channels; ++c) {
        Signal one;
        one.fs = signal.fs;
        one.channels = 1;
        one.samples = signal.samples;
        one.data.assign(signal.channel(c), signal.channel(c) + signal.samples);
        out_ << "CHANNEL\t" << c << std::endl;
        result.push_back(Decomposition::compute(settings, ws, one)[0]);
    }
    return result;
}

// src/mmp.cpp:23-45: decompose the channel sum, then refit each atom per channel
ChannelAtoms Mmp2Decomposition::compute(const DecompositionSettings& settings, Workspace* ws, const Signal& signal) {
    Signal sum;
    sum.fs = signal.fs;
    sum.channels = 1;
    sum.samples = signal.samples;
    sum.data.assign(signal.channel(0), signal.channel(0) + signal.samples);
    for (int c = 1; c < signal.channels; ++c)
        for (int n = 0; n < signal.samples; ++n) sum.data[n] += signal.channel(c)[n];
    std::vector<Atom> sumAtoms = Decomposition::compute(settings, ws, sum)[0];
    Signal residue = signal;
    ChannelAtoms result(static_cast<size_t>(signal.channels));
    for (int c = 0; c < signal.channels; ++c) {
        for (Atom atom : sumAtoms) {
            Workspace::subtractAtomFromSignal(atom, signal.fs, residue.channel(c), signal.samples, true);
            result[c].push_back(atom);
        }
    }
    return result;
}

// ---------------------------------------------------- BatchDecomposition

namespace {

std::vector<empi_item> items_of(const std::vector<Signal>& epochs, int variant, const DecompositionSettings& settings,
                                const empi_dict_params& dict) {
    std::vector<empi_item> items;
    for (const Signal& s : epochs) {
        empi_item it{};
        it.channels = s.channels;
        it.samples = s.samples;
        it.fs = s.fs;
        it.variant = variant;
        it.iterationMax = settings.iterationMax;
        it.energyPercent = settings.energyPercent;
        it.dict = dict;
        items.push_back(it);
    }
    return items;
}

bool same_items(const std::vector<empi_item>& x, const std::vector<empi_item>& y) {
    auto same = [](const empi_item& a, const empi_item& b) {
        return a.channels == b.channels && a.samples == b.samples && a.fs == b.fs && a.variant == b.variant &&
               a.iterationMax == b.iterationMax && a.energyPercent == b.energyPercent &&
               a.dict.energyError == b.dict.energyError && a.dict.scaleMin == b.dict.scaleMin &&
               a.dict.scaleMax == b.dict.scaleMax && a.dict.freqMax == b.dict.freqMax;
    };
    return x.size() == y.size() && std::equal(x.begin(), x.end(), y.begin(), same);
}

std::vector<float> samples_of(const std::vector<Signal>& epochs) {
    size_t total = 0;
    for (const Signal& s : epochs) total += s.data.size();
    std::vector<float> input;
    input.reserve(total);
    for (const Signal& s : epochs)
        for (double v : s.data) input.push_back(static_cast<float>(v));
    return input;
}

}  // namespace

void BatchDecomposition::prepare(const std::vector<Signal>& epochs, int variant, const DecompositionSettings& settings,
                                 const empi_dict_params& dict) {
    if (epochs.empty()) return;
    std::vector<empi_item> items = items_of(epochs, variant, settings, dict);
    if (plan_ && same_items(items, items_)) return;
    plan_.reset();
    items_.clear();
    empi_plan* created = nullptr;
    check(empi_plan_create(items.data(), static_cast<int>(items.size()), &created));
    plan_.reset(created);
    items_ = items;
}

bool BatchDecomposition::stage(const std::vector<Signal>& epochs, int variant, const DecompositionSettings& settings,
                               const empi_dict_params& dict) {
    if (!plan_ || epochs.empty() || !same_items(items_of(epochs, variant, settings, dict), items_)) return false;
    const std::vector<float> input = samples_of(epochs);
    check(empi_plan_stage(plan_.get(), input.data()));
    return true;
}

std::vector<BatchDecomposition::EpochResult> BatchDecomposition::run(const std::vector<Signal>& epochs, int variant,
                                                                     const DecompositionSettings& settings,
                                                                     const empi_dict_params& dict, bool staged) {
    std::vector<EpochResult> results(epochs.size());
    if (epochs.empty()) return results;
    if (staged) {
        check(empi_plan_run(plan_.get(), nullptr, 0, nullptr));
    } else {
        prepare(epochs, variant, settings, dict);
        const std::vector<float> input = samples_of(epochs);
        check(empi_plan_run(plan_.get(), input.data(), 0, nullptr));
    }
    empi_plan* plan = plan_.get();
    const int I = settings.iterationMax;
    for (size_t e = 0; e < epochs.size(); ++e) {
        const int C = epochs[e].channels;
        const int rows = variant == EMPI_SMP ? C : 1;
        std::vector<empi_atom> atoms(static_cast<size_t>(C) * I);
        std::vector<int> counts(static_cast<size_t>(C));
        std::vector<double> energies(static_cast<size_t>(rows) * (I + 1));
        EpochResult& r = results[e];
        r.status = empi_plan_item_result(plan, static_cast<int>(e), atoms.data(), counts.data(), energies.data());
        r.atomCount = empi_plan_item_atom_count(plan, static_cast<int>(e));
        r.counts = counts;
        r.atoms.resize(static_cast<size_t>(C));
        for (int c = 0; c < C; ++c)
            for (int i = 0; i < counts[c]; ++i) r.atoms[c].push_back(from_abi(atoms[static_cast<size_t>(c) * I + i]));
        for (int row = 0; row < rows; ++row)
            r.energies.emplace_back(energies.begin() + static_cast<long>(row) * (I + 1),
                                    energies.begin() + static_cast<long>(row + 1) * (I + 1));
    }
    return results;
}

void BatchDecomposition::print_progress(std::ostream& out, const std::vector<double>& energies, int count,
                                        size_t atomCount, int iterationMax, bool failed_search) {
    const double total = energies[0];
    const int lines = count + (failed_search ? 1 : 0);
    for (int i = 0; i < lines && i < iterationMax; ++i) {
        const double residue = energies[i];
        const double energyProgress = 100.0 * (1.0 - residue / total);
        const double totalProgress = std::max(100.0 * i / iterationMax, energyProgress);
        out << "ATOM\t" << i << '\t' << atomCount << '\t' << energyProgress << '\t' << totalProgress << std::endl;
    }
}

}  // namespace host
}  // namespace empi
