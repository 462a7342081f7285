#include "signal_reader.hpp"

#include <climits>

#include "config.hpp"

namespace empi {
namespace host {

SignalReader::SignalReader(const std::string& p) : path(p) {
    file_ = std::fopen(p.c_str(), "rb");
    if (!file_) throw Error("couldNotOpenSignalFile");
}

SignalReader::~SignalReader() {
    if (file_) std::fclose(file_);
}

Signal SignalReader::read_epoch(long long samples) {
    Signal s;
    s.fs = fs;
    s.channels = static_cast<int>(selectedChannels.size());
    std::vector<float> frame(static_cast<size_t>(channelCount));
    std::vector<std::vector<double>> cols(static_cast<size_t>(s.channels));
    long long stored = 0;
    while (samples-- > 0 &&
           std::fread(frame.data(), sizeof(float), frame.size(), file_) == frame.size()) {
        if (++stored > INT_MAX) throw Error("signalFileIsTooLongForThisMachine");
        for (int c = 0; c < s.channels; ++c) cols[c].push_back(frame[selectedChannels[c] - 1]);
    }
    s.samples = static_cast<int>(stored);
    s.data.resize(static_cast<size_t>(s.channels) * s.samples);
    for (int c = 0; c < s.channels; ++c)
        for (int n = 0; n < s.samples; ++n) s.data[static_cast<size_t>(c) * s.samples + n] = cols[c][n];
    return s;
}

void SignalReader::seek(long long sample) {
    if (std::fseek(file_, static_cast<long>(sizeof(float) * channelCount * sample), SEEK_SET))
        throw Error("signalSeekFailed");
}

SignalReaderForAllEpochs::SignalReaderForAllEpochs(const std::string& p, int epochSize)
    : SignalReader(p), epochSize_(epochSize) {}

Signal SignalReaderForAllEpochs::read() {
    Signal s = read_epoch(epochSize_);
    if (s.samples > 0 && s.samples != epochSize_) throw Error("fileIsTruncated");
    return s;
}

SignalReaderForSelectedEpochs::SignalReaderForSelectedEpochs(const std::string& p, int epochSize,
                                                             const std::vector<int>& epochs)
    : SignalReaderForAllEpochs(p, epochSize), epochs_(epochs.begin(), epochs.end()) {}

Signal SignalReaderForSelectedEpochs::read() {
    if (epochs_.empty()) return read_epoch(0);
    const int e = epochs_.front();
    epochs_.pop_front();
    seek(static_cast<long long>(epochSize_) * (e - 1));
    Signal s = read_epoch(epochSize_);
    if (s.samples != epochSize_) throw Error("fileIsTruncated");
    return s;
}

SignalReaderForWholeSignal::SignalReaderForWholeSignal(const std::string& p) : SignalReader(p) {}

Signal SignalReaderForWholeSignal::read() { return read_epoch(INT_MAX); }

}  // namespace host
}  // namespace empi
