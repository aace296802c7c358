// WAV decoding to mono f64 for the file entry points (audio.zig:41-130):
// RIFF/WAVE chunks, PCM 8/16/24/32-bit, IEEE float 32/64-bit, WAVE_FORMAT_
// EXTENSIBLE, channel mean; 16-bit PCM / 32768 as upstream.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace vq {

struct WavAudio {
    std::vector<double> samples;
    uint32_t sample_rate = 0;
};

// Decodes `data`; on failure returns false with the reference's error name.
bool decode_wav_mono(const uint8_t* data, size_t len, WavAudio& out, std::string& err);
bool load_wav_mono(const char* path, WavAudio& out, std::string& err);
// visqol.h code for an error name from the two functions above.
int wav_error_code(const std::string& err);

}  // namespace vq
