// Source data and audio I/O. See sources_io.hpp.
#include "sources_io.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace dg {

std::vector<std::vector<double>> parseCsvTable(const std::string& path, char sep) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open source table '" + path + "'");
    std::vector<std::vector<double>> rows;
    std::string line;
    int lineNo = 0;
    while (std::getline(in, line)) {
        ++lineNo;
        if (lineNo == 1 || line.empty() || line[0] == '#') continue;
        std::vector<double> row;
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, sep)) row.push_back(static_cast<double>(std::stof(cell)));
        rows.push_back(row);
    }
    return rows;
}

namespace {

template <typename T>
bool readLE(std::ifstream& in, T& v) {
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    return static_cast<bool>(in);
}

}  // namespace

std::vector<std::vector<double>> parseWavTable(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open WAV file '" + path + "'");
    char riff[4], wave[4];
    uint32_t riffSize;
    in.read(riff, 4);
    readLE(in, riffSize);
    in.read(wave, 4);
    if (!in || std::memcmp(riff, "RIFF", 4) != 0 || std::memcmp(wave, "WAVE", 4) != 0)
        throw std::runtime_error("'" + path + "' is not a RIFF/WAVE file");
    uint16_t format = 0, channels = 0, bits = 0;
    uint32_t rate = 0;
    std::vector<char> data;
    bool haveFmt = false, haveData = false;
    while (in && !(haveFmt && haveData)) {
        char id[4];
        uint32_t size;
        in.read(id, 4);
        if (!readLE(in, size)) break;
        if (std::memcmp(id, "fmt ", 4) == 0) {
            uint32_t byteRate;
            uint16_t blockAlign;
            readLE(in, format);
            readLE(in, channels);
            readLE(in, rate);
            readLE(in, byteRate);
            readLE(in, blockAlign);
            readLE(in, bits);
            if (size > 16) in.seekg(size - 16, std::ios::cur);
            haveFmt = true;
        } else if (std::memcmp(id, "data", 4) == 0) {
            data.resize(size);
            in.read(data.data(), size);
            haveData = true;
        } else {
            in.seekg(size + (size & 1), std::ios::cur);
        }
    }
    if (!haveFmt || !haveData || rate == 0) throw std::runtime_error("'" + path + "' has no PCM data");
    const int bytes = bits / 8;
    if (bytes < 1 || bytes > 4) throw std::runtime_error("'" + path + "': unsupported sample width");
    const size_t n = data.size() / bytes;
    std::vector<std::vector<double>> rows;
    rows.reserve(n);
    const double step = 1.0 / rate;
    double t = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&data[i * bytes]);
        float v = 0.0f;
        if (bits == 8) {
            v = static_cast<float>(static_cast<int8_t>(b[0])) / 127.0f;
        } else if (bits == 16) {
            int16_t s = static_cast<int16_t>(b[0] | (b[1] << 8));
            v = static_cast<float>(s) / 32767.0f;
        } else if (bits == 24) {
            int s = (b[2] & 0x80) ? static_cast<int>((0xffu << 24) | (b[2] << 16) | (b[1] << 8) | b[0])
                                  : ((b[2] << 16) | (b[1] << 8) | b[0]);
            v = static_cast<float>(s) / 8388607.0f;
        } else {
            int32_t s = static_cast<int32_t>(b[0] | (b[1] << 8) | (b[2] << 16) | (static_cast<uint32_t>(b[3]) << 24));
            v = static_cast<float>(s) / 2147483647.0f;
        }
        rows.push_back({t, static_cast<double>(v)});
        t += step;
    }
    return rows;
}

bool writeWav16(const std::string& path, const std::vector<float>& samples, uint32_t sampleRate) {
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    const uint32_t dataBytes = static_cast<uint32_t>(samples.size() * 2);
    auto put32 = [&](uint32_t v) { out.write(reinterpret_cast<const char*>(&v), 4); };
    auto put16 = [&](uint16_t v) { out.write(reinterpret_cast<const char*>(&v), 2); };
    out.write("RIFF", 4);
    put32(36 + dataBytes);
    out.write("WAVE", 4);
    out.write("fmt ", 4);
    put32(16);
    put16(1);
    put16(1);
    put32(sampleRate);
    put32(sampleRate * 2);
    put16(2);
    put16(16);
    out.write("data", 4);
    put32(dataBytes);
    for (float s : samples) {
        const int16_t v = static_cast<int16_t>(static_cast<int32_t>(s * 32767.0f));
        put16(static_cast<uint16_t>(v));
    }
    return static_cast<bool>(out);
}

double interpolateTable(const std::vector<std::vector<double>>& data, double t) {
    if (data.size() < 2) return 0.0;
    size_t lo = 0, hi = data.size() - 1;
    const double tmin = data[lo][0], tmax = data[hi][0];
    if (!(t >= tmin && t <= tmax)) return 0.0;
    while (hi > lo + 1) {
        const size_t m = (lo + hi) / 2;
        if (t < data[m][0])
            hi = m;
        else
            lo = m;
    }
    return data[lo][1] + ((data[hi][1] - data[lo][1]) / (data[hi][0] - data[lo][0])) * (t - data[lo][0]);
}

}  // namespace dg
