// Host-side codecs of the volume file formats (no device code).
//
// HxByteRLE, the run-length code of AmiraMesh byte lattices (read and written
// by biomedisa/features/amira_to_np/amira_data_stream.py:21-141): a stream of
// records, each a control byte c followed by
//   c & 0x80 set:   (c & 0x7f) literal bytes,
//   c & 0x80 clear: one byte repeated c times.
#include <string.h>

#include "bmx_internal.h"

extern "C" int64_t bmx_hxbyterle_decode(const uint8_t *in, int64_t n_in, uint8_t *out, int64_t n_out)
{
    BMX_CHECK((in || n_in == 0) && (out || n_out == 0) && n_in >= 0 && n_out >= 0,
              "bmx_hxbyterle_decode: bad arguments");
    int64_t i = 0, j = 0;
    while (i < n_in && j < n_out) {
        const uint8_t c = in[i++];
        const int64_t n = c & 0x7f;
        if (n == 0) continue;
        BMX_CHECK(j + n <= n_out, "bmx_hxbyterle_decode: stream longer than the lattice");
        if (c & 0x80) {
            BMX_CHECK(i + n <= n_in, "bmx_hxbyterle_decode: truncated literal record");
            memcpy(out + j, in + i, (size_t)n);
            i += n;
        } else {
            BMX_CHECK(i < n_in, "bmx_hxbyterle_decode: truncated run record");
            memset(out + j, in[i++], (size_t)n);
        }
        j += n;
    }
    BMX_CHECK(j == n_out, "bmx_hxbyterle_decode: stream shorter than the lattice");
    return j;
}

// Encoded size is at most n + ceil(n / 127); returns the bytes written.
extern "C" int64_t bmx_hxbyterle_encode(const uint8_t *in, int64_t n, uint8_t *out, int64_t cap)
{
    BMX_CHECK((in || n == 0) && (out || cap == 0) && n >= 0 && cap >= n + (n + 126) / 127,
              "bmx_hxbyterle_encode: bad arguments or output too small");
    int64_t i = 0, o = 0;
    while (i < n) {
        // a run of at least 3 equal bytes is coded as a run record
        int64_t r = 1;
        while (i + r < n && r < 127 && in[i + r] == in[i]) ++r;
        if (r >= 3) {
            out[o++] = (uint8_t)r;
            out[o++] = in[i];
            i += r;
            continue;
        }
        // otherwise literals up to the next run of 3
        int64_t l = 0;
        while (i + l < n && l < 127) {
            if (i + l + 2 < n && in[i + l] == in[i + l + 1] && in[i + l] == in[i + l + 2]) break;
            ++l;
        }
        out[o++] = (uint8_t)(0x80 | l);
        memcpy(out + o, in + i, (size_t)l);
        o += l;
        i += l;
    }
    return o;
}

// PNG scanline filters (ISO/IEC 15948, section 9): rows x (1 + rowbytes)
// filtered bytes -> rows x rowbytes raw bytes; bpp = bytes per complete pixel.
extern "C" int bmx_png_unfilter(const uint8_t *in, int64_t rows, int64_t rowbytes, int bpp, uint8_t *out)
{
    BMX_CHECK(in && out && rows >= 0 && rowbytes >= 0 && bpp >= 1 && bpp <= 8, "bmx_png_unfilter: bad arguments");
    for (int64_t r = 0; r < rows; ++r) {
        const uint8_t *src = in + r * (rowbytes + 1);
        const uint8_t ft = src[0];
        ++src;
        uint8_t *dst = out + r * rowbytes;
        const uint8_t *up = r > 0 ? dst - rowbytes : nullptr;
        for (int64_t i = 0; i < rowbytes; ++i) {
            const int a = i >= bpp ? dst[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= bpp) ? up[i - bpp] : 0;
            int pred;
            switch (ft) {
            case 0: pred = 0; break;
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (a + b) >> 1; break;
            case 4: {
                const int p = a + b - c;
                const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
                pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: return bmx_set_error(BMX_E_INVALID, "bmx_png_unfilter: filter type %d in row %lld", ft, (long long)r);
            }
            dst[i] = (uint8_t)(src[i] + pred);
        }
    }
    return BMX_OK;
}
