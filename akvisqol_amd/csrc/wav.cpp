#include "wav.h"

#include <string.h>

#include <fstream>
#include <iterator>

namespace vq {

namespace {
uint16_t u16(const uint8_t* p) { return (uint16_t)(p[0] | p[1] << 8); }
uint32_t u32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

constexpr uint16_t kPcm = 1, kFloat = 3, kExtensible = 0xFFFE;

// audio.zig:115-129
double sample_at(uint16_t fmt, uint16_t bits, const uint8_t* p) {
    if (fmt == kFloat) {
        if (bits == 32) {
            float f;
            const uint32_t b = u32(p);
            memcpy(&f, &b, 4);
            return (double)f;
        }
        uint64_t b = (uint64_t)u32(p) | (uint64_t)u32(p + 4) << 32;
        double d;
        memcpy(&d, &b, 8);
        return d;
    }
    switch (bits) {
        case 8: return ((double)p[0] - 128.0) / 128.0;
        case 16: return (double)(int16_t)u16(p) / 32768.0;
        case 24: {
            int32_t v = (int32_t)((uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16);
            if (v & 0x800000) v -= 0x1000000;
            return (double)v / 8388608.0;
        }
        default: return (double)(int32_t)u32(p) / 2147483648.0;
    }
}
}  // namespace

// audio.zig:48-113
bool decode_wav_mono(const uint8_t* d, size_t n, WavAudio& out, std::string& err) {
    if (n < 12 || memcmp(d, "RIFF", 4) || memcmp(d + 8, "WAVE", 4)) { err = "NotRiffWave"; return false; }
    uint16_t fmt = 0, nch = 0, bits = 0;
    uint32_t rate = 0;
    bool have_fmt = false;
    const uint8_t* pcm = nullptr;
    size_t pcm_len = 0;
    size_t off = 12;
    while (off + 8 <= n) {
        const uint8_t* id = d + off;
        const size_t size = u32(d + off + 4);
        const size_t body = off + 8;
        const size_t end = body + size < n ? body + size : n;  // oversized chunks are clamped
        const size_t blen = end - body;
        if (!memcmp(id, "fmt ", 4)) {
            if (blen < 16) { err = "UnsupportedFormat"; return false; }
            fmt = u16(d + body);
            nch = u16(d + body + 2);
            rate = u32(d + body + 4);
            bits = u16(d + body + 14);
            if (fmt == kExtensible && blen >= 26) fmt = u16(d + body + 24);
            have_fmt = true;
        } else if (!memcmp(id, "data", 4)) {
            pcm = d + body;
            pcm_len = blen;
            break;
        }
        off = body + size + (size & 1);
    }
    if (!have_fmt) { err = "MissingFmtChunk"; return false; }
    if (!pcm) { err = "MissingDataChunk"; return false; }
    if (nch == 0) { err = "ChannelCountZero"; return false; }
    const bool valid = fmt == kPcm ? (bits == 8 || bits == 16 || bits == 24 || bits == 32)
                                   : fmt == kFloat ? (bits == 32 || bits == 64) : false;
    if (!valid) { err = "UnsupportedFormat"; return false; }
    const size_t bps = bits / 8;
    const size_t frames = pcm_len / (bps * nch);
    if (frames == 0) { err = "EmptyAudio"; return false; }
    out.samples.resize(frames);
    out.sample_rate = rate;
    for (size_t f = 0; f < frames; f++) {
        double sum = 0.0;
        for (size_t c = 0; c < nch; c++) sum += sample_at(fmt, bits, pcm + (f * nch + c) * bps);
        out.samples[f] = nch == 1 ? sum : sum / (double)nch;
    }
    return true;
}

bool load_wav_mono(const char* path, WavAudio& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = "FileNotFound"; return false; }
    std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return decode_wav_mono(buf.data(), buf.size(), out, err);
}

// capi.zig:96-112: I/O and format errors are E_IO; signal-shape errors
// (empty audio, zero channels) are E_SIGNAL.
int wav_error_code(const std::string& err) {
    if (err == "EmptyAudio" || err == "ChannelCountZero") return -5;
    return -3;
}

}  // namespace vq
