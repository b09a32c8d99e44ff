// Source data and audio I/O of the front-end (host side, off the hot path).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dg {

// File sources (src/utils.cpp:138-177): ';'-separated rows, first line skipped, lines
// starting with '#' skipped, values rounded through float (the reference parses with stof).
// Empty lines are skipped (the reference would store an empty row and later read past it).
std::vector<std::vector<double>> parseCsvTable(const std::string& path, char sep = ';');

// WAV sources (src/utils.cpp:213-258 with the vendored reader src/wave/file.cc:231-289):
// every sample of every channel in file order at t = i / sample_rate; PCM 8/16/24/32-bit
// scaled by 1/INT_MAX of its width (int8 127, int16 32767, int24 8388607, int32 2147483647).
std::vector<std::vector<double>> parseWavTable(const std::string& path);

// 16-bit mono PCM writer (io::writeWave, src/utils.cpp:179-211 / src/wave/file.cc:291-352):
// sample -> int16 by truncation of sample*32767, sample rate = (uint32)(1/dt).
bool writeWav16(const std::string& path, const std::vector<float>& samples, uint32_t sampleRate);

// Linear interpolation in a (time, value) table by bisection (Sources::interpolate_value,
// include/configParser.h:38-58): 0 outside [t_first, t_last]. A table with fewer than two
// rows yields 0 (the reference loops forever on a single row).
double interpolateTable(const std::vector<std::vector<double>>& data, double t);

}  // namespace dg
