// Signal file readers (inc/io.hpp:127-172, src/io.cpp:9-85): float32
// samples in machine byte order, interleaved [t][channel]; selected channels
// are 1-based; epochs of a fixed size, a selected list of them, or the whole
// file as one epoch.
#pragma once

#include <cstdio>
#include <deque>
#include <string>
#include <vector>

#include "book.hpp"

namespace empi {
namespace host {

class SignalReader {
public:
    explicit SignalReader(const std::string& path);   // throws couldNotOpenSignalFile
    virtual ~SignalReader();
    virtual Signal read() = 0;                         // zero samples at the end

    std::string path;
    int channelCount = 0;
    double fs = 0;
    std::vector<int> selectedChannels;

protected:
    Signal read_epoch(long long samples);
    void seek(long long sample);

private:
    FILE* file_ = nullptr;
};

class SignalReaderForAllEpochs : public SignalReader {
public:
    SignalReaderForAllEpochs(const std::string& path, int epochSize);
    Signal read() override;

protected:
    int epochSize_;
};

class SignalReaderForSelectedEpochs : public SignalReaderForAllEpochs {
public:
    SignalReaderForSelectedEpochs(const std::string& path, int epochSize, const std::vector<int>& epochs);
    Signal read() override;

private:
    std::deque<int> epochs_;
};

class SignalReaderForWholeSignal : public SignalReader {
public:
    explicit SignalReaderForWholeSignal(const std::string& path);
    Signal read() override;
};

}  // namespace host
}  // namespace empi
