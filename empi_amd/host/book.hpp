// Signals, atoms and the MP5 "book" output format (inc/io.hpp:65-123,
// src/io.cpp:104-168): big-endian blocks, lengths back-patched, the original
// signal of every channel followed by its atoms as six float32 parameters.
#pragma once

#include <cstdio>
#include <string>
#include <vector>

namespace empi {
namespace host {

struct Atom {
    int type = 13;          // ATOM_GABOR (inc/base.hpp:22)
    double params[6] = {};  // modulus, amplitude, t [samples], s [samples], f [x Nyquist], phase
    int scale = -1, freq = -1, time = -1;
};

// One epoch of C channels, FP64 samples [c][n] (inc/base.hpp:35-67).
struct Signal {
    double fs = 0;
    int channels = 0;
    int samples = 0;
    std::vector<double> data;
    const double* channel(int c) const { return data.data() + static_cast<size_t>(c) * samples; }
    double* channel(int c) { return data.data() + static_cast<size_t>(c) * samples; }
    double energy() const;   // sum_c (sum_n x^2) / fs
};

typedef std::vector<std::vector<Atom>> ChannelAtoms;   // [channel][atom]

class BookWriter {
public:
    explicit BookWriter(const std::string& path);   // throws couldNotCreateOutputFile
    ~BookWriter();
    void write(int epoch, const Signal& signal, const ChannelAtoms& atoms);
    void close();

private:
    FILE* file_ = nullptr;
};

// Parsed book (for tests and tools): epochs of per-channel samples and atoms.
struct BookEpoch {
    int epoch = 0;
    int samples = 0;
    std::vector<std::vector<float>> signal;               // [channel][n]
    std::vector<std::vector<std::vector<float>>> atoms;   // [channel][atom][6]
};
struct Book {
    float fs = 0;
    int channels = 0;
    std::vector<BookEpoch> epochs;
};
Book read_book(const std::string& path);   // throws invalidBookFile

}  // namespace host
}  // namespace empi
