// Raw float64 output for the case programs: each field appended as nx*ny
// doubles, row-major, to the file named on the command line.
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "focalors/case_api.h"

struct RawOut {
    FILE* fp = nullptr;
    explicit RawOut(const char* path) : fp(std::fopen(path, "wb"))
    {
        if (!fp) throw std::runtime_error(std::string("cannot open ") + path);
    }
    ~RawOut() { std::fclose(fp); }
    void write(const double* p, size_t n)
    {
        if (std::fwrite(p, sizeof(double), n, fp) != n) throw std::runtime_error("short write");
    }
    void write(const field2& f) { write(f.data(), (size_t)f.get_size_n()); }
};

struct RawIn {
    std::vector<double> v;
    size_t              pos = 0;
    explicit RawIn(const char* path)
    {
        FILE* fp = std::fopen(path, "rb");
        if (!fp) throw std::runtime_error(std::string("cannot open ") + path);
        double x;
        while (std::fread(&x, sizeof(double), 1, fp) == 1) v.push_back(x);
        std::fclose(fp);
    }
    void read(double* p, size_t n)
    {
        if (pos + n > v.size()) throw std::runtime_error("input file too short");
        for (size_t k = 0; k < n; k++) p[k] = v[pos++];
    }
};
