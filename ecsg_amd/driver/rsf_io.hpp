// rsf_io.hpp -- minimal reader/writer of Madagascar RSF files for ecsg_run.
//
// An RSF dataset is a text header of key=value pairs (later assignments of
// a key win; values may be double-quoted) naming the binary data file with
// in=...; in="stdin" (or the control sequence 0x0C 0x0C 0x04 in the header
// file) means the data follows the header in the same file.  Supported data:
// esize=4 with data_format native_float, xdr_float (big endian) or
// ascii_float.  Written datasets are a header file plus "<header>@" holding
// native floats.
#pragma once
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace rsf {

inline std::string unquote(const std::string &v)
{
    if (v.size() >= 2 && (v.front() == '"' || v.front() == '\'') && v.back() == v.front()) return v.substr(1, v.size() - 2);
    return v;
}

// key=value pairs of a text (header lines, or a parameter file)
inline void parse_pairs(const std::string &text, std::map<std::string, std::string> &kv)
{
    size_t i = 0;
    const size_t n = text.size();
    while (i < n) {
        while (i < n && isspace((unsigned char)text[i])) i++;
        size_t k0 = i;
        while (i < n && !isspace((unsigned char)text[i]) && text[i] != '=') i++;
        if (i >= n || text[i] != '=') {
            while (i < n && !isspace((unsigned char)text[i])) i++;
            continue;
        }
        const std::string key = text.substr(k0, i - k0);
        i++;  // '='
        size_t v0 = i;
        if (i < n && (text[i] == '"' || text[i] == '\'')) {
            const char q = text[i++];
            while (i < n && text[i] != q) i++;
            if (i < n) i++;
        } else {
            while (i < n && !isspace((unsigned char)text[i])) i++;
        }
        if (!key.empty()) kv[key] = unquote(text.substr(v0, i - v0));
    }
}

struct Input {
    std::map<std::string, std::string> kv;
    std::vector<float> data;

    bool has(const std::string &k) const { return kv.count(k) != 0; }
    long geti(const std::string &k, long def) const { return has(k) ? std::stol(kv.at(k)) : def; }
    double getd(const std::string &k, double def) const { return has(k) ? std::stod(kv.at(k)) : def; }

    void open(const std::string &path)
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot open RSF header " + path);
        std::string all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        const char marker[3] = {'\014', '\014', '\004'};
        const size_t cut = all.find(std::string(marker, 3));
        const std::string header = all.substr(0, cut);
        parse_pairs(header, kv);
        long count = 1;
        for (int d = 1; d <= 9; d++) count *= geti("n" + std::to_string(d), 1);
        if (geti("esize", 4) != 4) throw std::runtime_error(path + ": only 4-byte elements are supported");
        const std::string fmt = has("data_format") ? kv.at("data_format") : "native_float";
        std::string bytes;
        const std::string in = has("in") ? kv.at("in") : "stdin";
        if (in == "stdin" || cut != std::string::npos) {
            if (cut == std::string::npos) throw std::runtime_error(path + ": in=stdin without appended data");
            bytes = all.substr(cut + 3);
        } else {
            std::ifstream d(in, std::ios::binary);
            if (!d) throw std::runtime_error(path + ": cannot open data file " + in);
            bytes.assign((std::istreambuf_iterator<char>(d)), std::istreambuf_iterator<char>());
        }
        data.resize((size_t)count);
        if (fmt == "ascii_float") {
            std::istringstream ss(bytes);
            for (long q = 0; q < count; q++)
                if (!(ss >> data[q])) throw std::runtime_error(path + ": short ascii data");
        } else {
            if ((long)bytes.size() < 4 * count) throw std::runtime_error(path + ": short binary data");
            const bool big = fmt == "xdr_float";
            if (!big && fmt != "native_float") throw std::runtime_error(path + ": unsupported data_format " + fmt);
            for (long q = 0; q < count; q++) {
                uint32_t u;
                memcpy(&u, bytes.data() + 4 * q, 4);
                if (big) u = __builtin_bswap32(u);
                memcpy(&data[q], &u, 4);
            }
        }
    }
};

struct Output {
    std::string path;
    std::map<std::string, std::string> kv;
    std::vector<std::string> order;
    std::ofstream data;

    void put(const std::string &k, const std::string &v)
    {
        if (!kv.count(k)) order.push_back(k);
        kv[k] = v;
    }
    void put(const std::string &k, long v) { put(k, std::to_string(v)); }
    void put(const std::string &k, int v) { put(k, std::to_string(v)); }
    // float-valued header entry (the value as a float, as RSF stores it)
    void putf(const std::string &k, double v)
    {
        char b[64];
        snprintf(b, sizeof b, "%.9g", (double)(float)v);
        put(k, std::string(b));
    }
    void putq(const std::string &k, const std::string &v) { put(k, "\"" + v + "\""); }

    // writes the header at path and opens path + "@" for the native floats
    void open()
    {
        const std::string dpath = path + "@";
        std::ofstream h(path);
        if (!h) throw std::runtime_error("cannot write " + path);
        h << "ecsg_run\n";
        for (const auto &k : order) h << "\t" << k << "=" << kv[k] << "\n";
        h << "\tesize=4\n\tdata_format=\"native_float\"\n\tin=\"" << dpath << "\"\n";
        data.open(dpath, std::ios::binary | std::ios::trunc);
        if (!data) throw std::runtime_error("cannot write " + dpath);
    }
    void append(const float *v, size_t n)
    {
        data.write(reinterpret_cast<const char *>(v), (std::streamsize)(4 * n));
        if (!data) throw std::runtime_error("write failed on " + path + "@");
    }
    void close()
    {
        if (data.is_open()) data.close();
    }
    void write(const std::vector<float> &v)
    {
        open();
        append(v.data(), v.size());
        close();
    }
};

}  // namespace rsf
