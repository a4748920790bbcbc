// Raw / NRRD image I/O (see image_io.hpp).
#include "image_io.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace lc2io {

namespace {

size_t dtype_bytes(lc2e_dtype t) {
    switch (t) {
        case LC2E_U8: return 1;
        case LC2E_I16: return 2;
        case LC2E_F16: return 2;
        case LC2E_F32: return 4;
        default: throw std::runtime_error("image: unsupported dtype");
    }
}

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

}  // namespace

lc2e_dtype parse_dtype(const std::string& name) {
    const std::string n = lower(trim(name));
    if (n == "float32" || n == "float" || n == "f32" || n == "2d") return LC2E_F32;
    if (n == "float16" || n == "half" || n == "f16") return LC2E_F16;
    if (n == "int16" || n == "short" || n == "i16" || n == "signed short") return LC2E_I16;
    if (n == "uint8" || n == "uchar" || n == "u8" || n == "unsigned char") return LC2E_U8;
    throw std::runtime_error("image: unknown data type '" + name + "'");
}

Image read_raw(const std::string& path, const int dims[3], const float spacing[3],
               const float origin[3], lc2e_dtype dtype) {
    Image im;
    im.dtype = dtype;
    for (int a = 0; a < 3; ++a) {
        im.geom.dims[a] = dims[a] > 0 ? dims[a] : 1;
        im.geom.spacing[a] = spacing[a];
        im.geom.origin[a] = origin[a];
    }
    const size_t n = (size_t)im.geom.dims[0] * im.geom.dims[1] * im.geom.dims[2];
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("image: cannot open " + path);
    im.bytes.resize(n * dtype_bytes(dtype));
    f.read(reinterpret_cast<char*>(im.bytes.data()), (std::streamsize)im.bytes.size());
    if ((size_t)f.gcount() != im.bytes.size())
        throw std::runtime_error("image: " + path + " holds fewer voxels than its dimensions");
    return im;
}

Image read_nrrd(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("image: cannot open " + path);
    std::string line;
    std::getline(f, line);
    if (line.rfind("NRRD", 0) != 0) throw std::runtime_error("image: " + path + " is not NRRD");
    Image im;
    int dim = 0;
    int sizes[3] = {1, 1, 1};
    float sp[3] = {1, 1, 1}, org[3] = {0, 0, 0};
    std::string encoding = "raw", endian = "little", data_file;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) break;
        if (line[0] == '#') continue;
        const size_t c = line.find(':');
        if (c == std::string::npos) continue;
        const std::string key = lower(trim(line.substr(0, c)));
        std::string val = trim(line.substr(c + 1));
        if (!val.empty() && val[0] == '=') val = trim(val.substr(1));
        if (key == "type") im.dtype = parse_dtype(val);
        else if (key == "dimension") dim = std::stoi(val);
        else if (key == "sizes") {
            std::istringstream is(val);
            for (int a = 0; a < 3 && is >> sizes[a]; ++a) {}
        } else if (key == "spacings") {
            std::istringstream is(val);
            for (int a = 0; a < 3 && is >> sp[a]; ++a) {}
        } else if (key == "space directions") {
            // axis-aligned directions: the diagonal entries are the spacings
            int a = 0;
            size_t p = 0;
            while (a < 3 && (p = val.find('(', p)) != std::string::npos) {
                const size_t q = val.find(')', p);
                std::string inner = val.substr(p + 1, q - p - 1);
                for (auto& ch : inner)
                    if (ch == ',') ch = ' ';
                std::istringstream is(inner);
                float v[3] = {0, 0, 0};
                for (int k = 0; k < 3 && is >> v[k]; ++k) {}
                sp[a] = v[a] != 0.0f ? std::fabs(v[a]) : 1.0f;
                ++a;
                p = q;
            }
        } else if (key == "space origin") {
            std::string inner = val;
            for (auto& ch : inner)
                if (ch == '(' || ch == ')' || ch == ',') ch = ' ';
            std::istringstream is(inner);
            for (int a = 0; a < 3 && is >> org[a]; ++a) {}
        } else if (key == "encoding") encoding = lower(val);
        else if (key == "endian") endian = lower(val);
        else if (key == "data file" || key == "datafile") data_file = val;
    }
    if (dim < 1 || dim > 3) throw std::runtime_error("image: NRRD dimension must be 1..3");
    if (encoding != "raw") throw std::runtime_error("image: only raw NRRD encoding is supported");
    if (endian != "little" && dtype_bytes(im.dtype) > 1)
        throw std::runtime_error("image: big-endian NRRD is not supported");
    for (int a = 0; a < 3; ++a) {
        im.geom.dims[a] = a < dim ? sizes[a] : 1;
        im.geom.spacing[a] = sp[a];
        im.geom.origin[a] = org[a];
    }
    const size_t n = (size_t)im.geom.dims[0] * im.geom.dims[1] * im.geom.dims[2];
    im.bytes.resize(n * dtype_bytes(im.dtype));
    if (!data_file.empty()) {
        std::string dp = data_file;
        const size_t slash = path.find_last_of('/');
        if (dp[0] != '/' && slash != std::string::npos) dp = path.substr(0, slash + 1) + dp;
        std::ifstream df(dp, std::ios::binary);
        if (!df) throw std::runtime_error("image: cannot open NRRD data file " + dp);
        df.read(reinterpret_cast<char*>(im.bytes.data()), (std::streamsize)im.bytes.size());
        if ((size_t)df.gcount() != im.bytes.size()) throw std::runtime_error("image: short NRRD data file");
    } else {
        f.read(reinterpret_cast<char*>(im.bytes.data()), (std::streamsize)im.bytes.size());
        if ((size_t)f.gcount() != im.bytes.size()) throw std::runtime_error("image: short NRRD payload");
    }
    return im;
}

Image read_image(const std::string& path, const int dims[3], const float spacing[3],
                 const float origin[3], lc2e_dtype dtype) {
    const std::string l = lower(path);
    const bool nrrd = l.size() > 5 && (l.substr(l.size() - 5) == ".nrrd" || l.substr(l.size() - 5) == ".nhdr");
    return nrrd ? read_nrrd(path) : read_raw(path, dims, spacing, origin, dtype);
}

void write_nrrd(const std::string& path, const float* data, const lc2e_geom& g) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("image: cannot write " + path);
    f << "NRRD0004\n"
      << "type: float\n"
      << "dimension: 3\n"
      << "space: scanner-xyz\n"
      << "sizes: " << g.dims[0] << " " << g.dims[1] << " " << g.dims[2] << "\n"
      << "space directions: (" << g.spacing[0] << ",0,0) (0," << g.spacing[1] << ",0) (0,0,"
      << g.spacing[2] << ")\n"
      << "kinds: domain domain domain\n"
      << "endian: little\n"
      << "encoding: raw\n"
      << "space origin: (" << g.origin[0] << "," << g.origin[1] << "," << g.origin[2] << ")\n\n";
    const size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(float)));
}

void write_raw(const std::string& path, const void* data, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("image: cannot write " + path);
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
}

}  // namespace lc2io
