// NumPy .npy v1.0 writer (host code): the file the reference's field export
// writes, byte for byte (src/core/output/npy_writer.cpp:33-182): the magic
// "\x93NUMPY", version 1.0, a little-endian u16 header length, the header
// dict "{'descr': '<f4', 'fortran_order': False, 'shape': (a, b, c), }" padded
// with spaces so that the preamble, dict and newline end on a 16-byte
// boundary, then the C-order payload.
#include "nimbus_hip.h"
#include "nimbus_internal.hpp"

#include <cstdio>
#include <string>

namespace nimbus
{

int write_npy(const char* path, const void* data, int dtype, int ndim, const std::int64_t* shape)
{
    if (!path || !data || !shape || ndim < 1 || ndim > 3) return set_error(NIMBUS_EINVAL, "npy: invalid argument");
    if (dtype != NIMBUS_NPY_F4 && dtype != NIMBUS_NPY_F8) return set_error(NIMBUS_EINVAL, "npy: dtype is f4 or f8");
    std::string dict = std::string("{'descr': '") + (dtype == NIMBUS_NPY_F4 ? "<f4" : "<f8") +
                       "', 'fortran_order': False, 'shape': (";
    std::size_t count = 1;
    for (int d = 0; d < ndim; ++d)
    {
        if (shape[d] <= 0 || shape[d] > 0x7fffffff) return set_error(NIMBUS_EINVAL, "npy: extents must be positive");
        dict += (d ? ", " : "") + std::to_string(shape[d]);
        count *= std::size_t(shape[d]);
    }
    dict += ndim == 1 ? ",), }" : "), }";
    const std::size_t preamble = 10;                       // magic, version, header length
    const std::size_t pad = (16 - (preamble + dict.size() + 1) % 16) % 16;
    const std::size_t hlen = dict.size() + pad + 1;
    if (hlen > 65535) return set_error(NIMBUS_EINVAL, "npy: header too long");
    std::string head("\x93NUMPY\x01\x00", 8);
    head += char(hlen & 0xff);
    head += char((hlen >> 8) & 0xff);
    head += dict;
    head.append(pad, ' ');
    head += '\n';
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(NIMBUS_EINVAL, std::string("npy: cannot open ") + path);
    const std::size_t bytes = count * (dtype == NIMBUS_NPY_F4 ? 4 : 8);
    const bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
                    std::fwrite(data, 1, bytes, f) == bytes;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return set_error(NIMBUS_EINVAL, std::string("npy: write failed: ") + path);
    return NIMBUS_OK;
}

}  // namespace nimbus

extern "C" int nimbus_npy_write(const char* path, const void* data, int dtype, int ndim, const int64_t* shape)
{
    return nimbus::write_npy(path, data, dtype, ndim, shape);
}
