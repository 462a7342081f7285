#include "book.hpp"

#include <cstdint>
#include <cstring>

#include "config.hpp"

namespace empi {
namespace host {

double Signal::energy() const {
    double total = 0.0;
    for (int c = 0; c < channels; ++c) {
        double sum = 0.0;
        const double* x = channel(c);
        for (int n = 0; n < samples; ++n) sum += x[n] * x[n];
        total += sum / fs;
    }
    return total;
}

namespace {

// big-endian byte writers
void be_u8(std::vector<unsigned char>& b, unsigned v) { b.push_back(static_cast<unsigned char>(v)); }
void be_u16(std::vector<unsigned char>& b, unsigned v) {
    b.push_back(static_cast<unsigned char>((v >> 8) & 0xFF));
    b.push_back(static_cast<unsigned char>(v & 0xFF));
}
void be_u32(std::vector<unsigned char>& b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back(static_cast<unsigned char>((v >> s) & 0xFF));
}
void be_f32(std::vector<unsigned char>& b, double v) {
    const float f = static_cast<float>(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    be_u32(b, u);
}
void patch_u32(std::vector<unsigned char>& b, size_t at, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[at + i] = static_cast<unsigned char>((v >> (24 - 8 * i)) & 0xFF);
}

}  // namespace

BookWriter::BookWriter(const std::string& path) {
    file_ = std::fopen(path.c_str(), "wb");
    if (!file_) throw Error("couldNotCreateOutputFile");
}

BookWriter::~BookWriter() { close(); }

void BookWriter::close() {
    if (file_) {
        std::fclose(file_);
        file_ = nullptr;
    }
}

void BookWriter::write(int epoch, const Signal& signal, const ChannelAtoms& atoms) {
    std::vector<unsigned char> b;
    const int C = signal.channels, N = signal.samples;
    if (epoch == 1) {
        // header (src/io.cpp:113-123; its fixed values are the reference's)
        const char magic[6] = {'M', 'P', 'v', '5', '.', '0'};
        b.insert(b.end(), magic, magic + 6);
        be_u8(b, 1);
        be_u32(b, 4);
        const char comment[4] = {'e', 'm', 'p', 'i'};
        b.insert(b.end(), comment, comment + 4);
        be_u8(b, 2);
        be_u32(b, 27);
        be_u8(b, 5);
        be_u8(b, 10);
        be_f32(b, C ? signal.fs : 0.0);
        be_f32(b, 1.0);   // points per microvolt
        be_u16(b, static_cast<unsigned>(C));
        be_u8(b, 6);
        be_u8(b, 13);
        be_f32(b, 100.0);   // energy percent
        be_u32(b, 1);       // iteration count
        be_u32(b, 0);       // dictionary size
        be_u8(b, 'F');
    }
    const size_t data_at = b.size();
    be_u8(b, 7);
    be_u32(b, 0);
    be_u16(b, static_cast<unsigned>(epoch));
    be_u32(b, static_cast<uint32_t>(N));
    for (int c = 0; c < C; ++c) {
        be_u8(b, 8);
        be_u32(b, static_cast<uint32_t>(2 + 4 * N));
        be_u16(b, static_cast<unsigned>(c + 1));
        const double* x = signal.channel(c);
        for (int n = 0; n < N; ++n) be_f32(b, x[n]);
        const size_t atoms_at = b.size();
        be_u8(b, 9);
        be_u32(b, 0);
        be_u16(b, static_cast<unsigned>(c + 1));
        if (c < static_cast<int>(atoms.size())) {
            for (const Atom& a : atoms[c]) {
                be_u8(b, static_cast<unsigned>(a.type));
                be_u8(b, 24);
                for (int p = 0; p < 6; ++p) be_f32(b, a.params[p]);
            }
        }
        patch_u32(b, atoms_at + 1, static_cast<uint32_t>(b.size() - atoms_at - 5));
    }
    patch_u32(b, data_at + 1, static_cast<uint32_t>(b.size() - data_at - 5));
    if (std::fwrite(b.data(), 1, b.size(), file_) != b.size()) throw Error("couldNotWriteOutputFile");
}

namespace {

struct Cursor {
    const std::vector<unsigned char>& b;
    size_t at = 0;
    void need(size_t n) const {
        if (at + n > b.size()) throw Error("invalidBookFile");
    }
    unsigned u8() {
        need(1);
        return b[at++];
    }
    unsigned u16() {
        need(2);
        unsigned v = (b[at] << 8) | b[at + 1];
        at += 2;
        return v;
    }
    uint32_t u32() {
        need(4);
        uint32_t v = (static_cast<uint32_t>(b[at]) << 24) | (static_cast<uint32_t>(b[at + 1]) << 16) |
                     (static_cast<uint32_t>(b[at + 2]) << 8) | b[at + 3];
        at += 4;
        return v;
    }
    float f32() {
        uint32_t u = u32();
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
};

}  // namespace

Book read_book(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error("couldNotOpenBookFile");
    std::vector<unsigned char> bytes;
    unsigned char tmp[65536];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) bytes.insert(bytes.end(), tmp, tmp + n);
    std::fclose(f);
    Cursor c{bytes};
    c.need(6);
    if (std::memcmp(bytes.data(), "MPv5.0", 6) != 0) throw Error("invalidBookFile");
    c.at = 6;
    Book book;
    // header blocks until the first epoch block (id 7)
    while (c.at < bytes.size() && bytes[c.at] != 7) {
        const unsigned id = c.u8();
        const uint32_t len = c.u32();
        const size_t end = c.at + len;
        if (id == 2) {
            while (c.at < end) {
                const unsigned sid = c.u8();
                const unsigned slen = c.u8();
                const size_t send = c.at + slen;
                if (sid == 5) {
                    book.fs = c.f32();
                    c.f32();
                    book.channels = static_cast<int>(c.u16());
                }
                c.at = send;
            }
        }
        c.at = end;
    }
    while (c.at < bytes.size()) {
        if (c.u8() != 7) throw Error("invalidBookFile");
        const uint32_t len = c.u32();
        const size_t end = c.at + len;
        BookEpoch e;
        e.epoch = static_cast<int>(c.u16());
        e.samples = static_cast<int>(c.u32());
        while (c.at < end) {
            const unsigned id = c.u8();
            const uint32_t blen = c.u32();
            const size_t bend = c.at + blen;
            const unsigned ch = c.u16();
            if (ch < 1) throw Error("invalidBookFile");
            if (e.signal.size() < ch) e.signal.resize(ch);
            if (e.atoms.size() < ch) e.atoms.resize(ch);
            if (id == 8) {
                while (c.at < bend) e.signal[ch - 1].push_back(c.f32());
            } else if (id == 9) {
                while (c.at < bend) {
                    c.u8();
                    const unsigned alen = c.u8();
                    std::vector<float> p;
                    for (unsigned i = 0; i < alen / 4; ++i) p.push_back(c.f32());
                    e.atoms[ch - 1].push_back(p);
                }
            }
            c.at = bend;
        }
        book.epochs.push_back(e);
    }
    return book;
}

}  // namespace host
}  // namespace empi
