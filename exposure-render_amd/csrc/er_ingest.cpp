/*
 * er_ingest.cpp - volume ingestion: MetaImage (.mhd header + raw data) and headerless raw
 * files into the uint16 voxel layout er_bind_volume takes (x fastest, then y, then z).
 *
 * The reference leaves loading to its (absent) VTK/Qt application, which read MetaImage
 * volumes such as manix_small.mhd; ErVolume::BindVoxels then takes unsigned shorts
 * (Source/ervolume.h). Here the file is read on the host in one pass; the device-side
 * structures (cell records, macrocell grids) are built by er_bind_volume as for any voxels.
 *
 * Element types map to uint16 as follows: unsigned 8- and 16-bit values are taken as they
 * are; other integer types are shifted by their minimum (so CT Hounsfield units, -1024 and
 * up, start at 0), floating-point values are rescaled from [min, max] to [0, 65535]. Both
 * are reported as u16 = (value - offset) * scale, so that transfer-function node positions
 * can be mapped from the file's units. Values beyond 65535 after the shift are clamped.
 */
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/er_capi.h"

extern "C" int er_set_error_message(int status, const char* message);

namespace {

int fail(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return er_set_error_message(ER_ERROR, buf);
}

struct TypeInfo {
    const char* name;
    int type;
    int bytes;
};

const TypeInfo TYPES[] = {
    {"MET_UCHAR", ER_ELEMENT_UINT8, 1},   {"MET_CHAR", ER_ELEMENT_INT8, 1},     {"MET_USHORT", ER_ELEMENT_UINT16, 2},
    {"MET_SHORT", ER_ELEMENT_INT16, 2},   {"MET_UINT", ER_ELEMENT_UINT32, 4},   {"MET_INT", ER_ELEMENT_INT32, 4},
    {"MET_FLOAT", ER_ELEMENT_FLOAT32, 4}, {"MET_DOUBLE", ER_ELEMENT_FLOAT64, 8},
};

int type_bytes(int type)
{
    for (const TypeInfo& t : TYPES)
        if (t.type == type) return t.bytes;
    return 0;
}

struct Header {
    int dims[3] = {0, 0, 0};
    float spacing[3] = {1.0f, 1.0f, 1.0f};
    int type = -1;
    int msb = 0;
    long long header_size = 0;   /* -1: the data is the last bytes of the file */
    std::string data_file;       /* resolved path; empty with LOCAL */
    bool local = false;
    size_t data_offset = 0;      /* LOCAL: where the header ends */
};

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool iequals(const std::string& a, const char* b)
{
    const size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

/* MetaImage text header: `Key = Value` lines up to ElementDataFile, which must come last */
int parse_mhd(const char* path, Header& h)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail("cannot open %s", path);
    std::string line;
    int ndims = 0, channels = 1;
    bool have_file = false;
    bool compressed = false;
    int c;
    size_t pos = 0;
    while (!have_file) {
        line.clear();
        while ((c = std::fgetc(f)) != EOF && c != '\n') line.push_back((char)c);
        pos += line.size() + (c == '\n' ? 1 : 0);
        if (c == EOF && line.empty()) break;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (iequals(key, "NDims")) {
            ndims = std::atoi(val.c_str());
        } else if (iequals(key, "DimSize")) {
            std::sscanf(val.c_str(), "%d %d %d", &h.dims[0], &h.dims[1], &h.dims[2]);
        } else if (iequals(key, "ElementSpacing") || iequals(key, "ElementSize")) {
            std::sscanf(val.c_str(), "%f %f %f", &h.spacing[0], &h.spacing[1], &h.spacing[2]);
        } else if (iequals(key, "ElementType")) {
            for (const TypeInfo& t : TYPES)
                if (iequals(val, t.name)) h.type = t.type;
            if (h.type < 0) { std::fclose(f); return fail("%s: unsupported ElementType %s", path, val.c_str()); }
        } else if (iequals(key, "BinaryDataByteOrderMSB") || iequals(key, "ElementByteOrderMSB")) {
            h.msb = iequals(val, "True") || val == "1";
        } else if (iequals(key, "HeaderSize")) {
            h.header_size = std::atoll(val.c_str());
        } else if (iequals(key, "CompressedData")) {
            compressed = iequals(val, "True") || val == "1";
        } else if (iequals(key, "ElementNumberOfChannels")) {
            channels = std::atoi(val.c_str());
        } else if (iequals(key, "ElementDataFile")) {
            have_file = true;
            if (iequals(val, "LOCAL")) {
                h.local = true;
                h.data_offset = pos;
            } else if (iequals(val, "LIST") || val.find('%') != std::string::npos) {
                std::fclose(f);
                return fail("%s: ElementDataFile lists are not supported", path);
            } else {
                std::string dir(path);
                const size_t slash = dir.find_last_of('/');
                h.data_file = (val[0] == '/' || slash == std::string::npos) ? val : dir.substr(0, slash + 1) + val;
            }
        }
    }
    std::fclose(f);
    if (!have_file) return fail("%s: no ElementDataFile", path);
    if (ndims != 3) return fail("%s: NDims must be 3 (got %d)", path, ndims);
    if (channels != 1) return fail("%s: %d channels per element; one is supported", path, channels);
    if (compressed) return fail("%s: compressed data is not supported", path);
    if (h.type < 0) return fail("%s: no ElementType", path);
    for (int i = 0; i < 3; ++i)
        if (h.dims[i] <= 0) return fail("%s: invalid DimSize", path);
    return ER_OK;
}

/* the raw bytes of dims[0] * dims[1] * dims[2] elements */
int read_data(const char* path, long long header_size, size_t data_offset, size_t bytes, std::vector<unsigned char>& out)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail("cannot open %s", path);
    if (std::fseek(f, 0, SEEK_END) != 0) { std::fclose(f); return fail("%s: cannot seek", path); }
    const long long size = std::ftell(f);
    long long start = (long long)data_offset + std::max(0LL, header_size);
    if (header_size < 0) start = size - (long long)bytes;
    if (start < 0 || start + (long long)bytes > size) {
        std::fclose(f);
        return fail("%s: %lld bytes of data expected from offset %lld, the file has %lld", path, (long long)bytes, start, size);
    }
    out.resize(bytes);
    if (std::fseek(f, (long)start, SEEK_SET) != 0 || std::fread(out.data(), 1, bytes, f) != bytes) {
        std::fclose(f);
        return fail("%s: short read", path);
    }
    std::fclose(f);
    return ER_OK;
}

double element(const unsigned char* p, int type, int msb)
{
    unsigned char b[8];
    const int n = type_bytes(type);
    for (int i = 0; i < n; ++i) b[i] = msb ? p[n - 1 - i] : p[i];   /* to little endian */
    switch (type) {
        case ER_ELEMENT_UINT8: return b[0];
        case ER_ELEMENT_INT8: return (signed char)b[0];
        case ER_ELEMENT_UINT16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case ER_ELEMENT_INT16: { int16_t v; std::memcpy(&v, b, 2); return v; }
        case ER_ELEMENT_UINT32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        case ER_ELEMENT_INT32: { int32_t v; std::memcpy(&v, b, 4); return v; }
        case ER_ELEMENT_FLOAT32: { float v; std::memcpy(&v, b, 4); return v; }
        case ER_ELEMENT_FLOAT64: { double v; std::memcpy(&v, b, 8); return v; }
    }
    return 0.0;
}

/* convert n elements to uint16 (the mapping in the file comment) */
int convert(const std::vector<unsigned char>& raw, int type, int msb, size_t n, uint16_t* dst, float* offset, float* scale)
{
    const int nb = type_bytes(type);
    double off = 0.0, sc = 1.0;
    if (type != ER_ELEMENT_UINT8 && type != ER_ELEMENT_UINT16) {
        double lo = INFINITY, hi = -INFINITY;
        for (size_t i = 0; i < n; ++i) {
            const double v = element(&raw[i * nb], type, msb);
            if (std::isnan(v)) return fail("the volume holds NaN values");
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
        off = lo;
        if (type == ER_ELEMENT_FLOAT32 || type == ER_ELEMENT_FLOAT64) sc = hi > lo ? 65535.0 / (hi - lo) : 1.0;
    }
    for (size_t i = 0; i < n; ++i) {
        double v = (element(&raw[i * nb], type, msb) - off) * sc;
        if (type == ER_ELEMENT_FLOAT32 || type == ER_ELEMENT_FLOAT64) v = std::floor(v + 0.5);
        dst[i] = (uint16_t)std::min(65535.0, std::max(0.0, v));
    }
    if (offset) *offset = (float)off;
    if (scale) *scale = (float)sc;
    return ER_OK;
}

} /* namespace */

extern "C" {

ER_API int er_mhd_info(const char* path, int resolution[3], float spacing[3], int* element_type)
{
    if (!path) return fail("path is NULL");
    Header h;
    int st = parse_mhd(path, h);
    if (st) return st;
    for (int i = 0; i < 3; ++i) {
        if (resolution) resolution[i] = h.dims[i];
        if (spacing) spacing[i] = h.spacing[i];
    }
    if (element_type) *element_type = h.type;
    return ER_OK;
}

ER_API int er_mhd_read(const char* path, uint16_t* voxels, size_t count, float* offset, float* scale)
{
    if (!path || !voxels) return fail("NULL argument");
    Header h;
    int st = parse_mhd(path, h);
    if (st) return st;
    const size_t n = (size_t)h.dims[0] * h.dims[1] * h.dims[2];
    if (count < n) return fail("need room for %zu voxels, got %zu", n, count);
    std::vector<unsigned char> raw;
    const size_t bytes = n * type_bytes(h.type);
    if (h.local) st = read_data(path, h.header_size > 0 ? h.header_size : 0, h.data_offset, bytes, raw);
    else st = read_data(h.data_file.c_str(), h.header_size, 0, bytes, raw);
    if (st) return st;
    return convert(raw, h.type, h.msb, n, voxels, offset, scale);
}

ER_API int er_raw_read(const char* path, const int resolution[3], int element_type, int msb, long long header_bytes,
                       uint16_t* voxels, size_t count, float* offset, float* scale)
{
    if (!path || !resolution || !voxels) return fail("NULL argument");
    if (type_bytes(element_type) == 0) return fail("unknown element type %d", element_type);
    for (int i = 0; i < 3; ++i)
        if (resolution[i] <= 0) return fail("invalid resolution");
    const size_t n = (size_t)resolution[0] * resolution[1] * resolution[2];
    if (count < n) return fail("need room for %zu voxels, got %zu", n, count);
    std::vector<unsigned char> raw;
    const int st = read_data(path, header_bytes, 0, n * type_bytes(element_type), raw);
    if (st) return st;
    return convert(raw, element_type, msb, n, voxels, offset, scale);
}

} /* extern "C" */
