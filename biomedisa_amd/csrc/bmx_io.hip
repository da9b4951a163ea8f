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
