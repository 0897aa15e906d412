// PNG, JPEG and binary PNM decoders producing RGBA8 for diffuse textures.
//
// Byte-for-byte agreement with the reference's loader (stb_image 2.30,
// stbi_load(..., 4), called from Src/Core/texture.cu:207-214) fixes these
// choices, each written here from its format specification:
//  - PNG: 16-bit samples keep their high byte; 1/2/4-bit grey is scaled by
//    255 / (2^depth - 1); a tRNS key makes matching pixels transparent;
//    grey expands to (g, g, g, a).
//  - JPEG: the integer LLM inverse DCT (Loeffler-Ligtenberg-Moschytz, as in the
//    IJG "islow" method) with 12-bit fixed-point multipliers truncated as
//    (int)(c * 4096 + 0.5), two extra bits kept between the passes; triangle
//    ("fancy") chroma upsampling for 2x1, 1x2 and 2x2 subsampling and
//    nearest-neighbour otherwise; YCbCr -> RGB in 20-bit fixed point with the
//    Cb term of green truncated to 16 fractional bits.
//  - PNM: raw samples, no maxval rescaling; of a 16-bit sample the second
//    (least significant) byte, as a little-endian host reads it.
#include "image.h"

#include <zlib.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace rt {
namespace {

[[noreturn]] void fail(const std::string& what) { throw std::runtime_error(what); }

constexpr uint64_t kMaxSide = 1u << 24;
constexpr uint64_t kMaxPixels = 1ull << 28;

void check_dims(uint64_t w, uint64_t h, const char* fmt) {
    if (w == 0 || h == 0) fail(std::string(fmt) + ": zero-sized image");
    if (w > kMaxSide || h > kMaxSide || w * h > kMaxPixels) fail(std::string(fmt) + ": image too large");
}

uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}
uint32_t be16(const uint8_t* p) { return (uint32_t)p[0] << 8 | (uint32_t)p[1]; }

void put_px(uint8_t* o, int r, int g, int b, int a) {
    o[0] = (uint8_t)r;
    o[1] = (uint8_t)g;
    o[2] = (uint8_t)b;
    o[3] = (uint8_t)a;
}

// =============================================================== PNG
int paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

struct PngHeader {
    uint32_t w = 0, h = 0;
    int depth = 0, color = -1, interlace = 0, channels = 0;
};

bool png_depth_ok(int color, int depth) {
    switch (color) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2: case 4: case 6: return depth == 8 || depth == 16;
        default: return false;
    }
}

// Reverse the scanline filters of one (sub)image in place: rows of `stride`
// bytes, each preceded by its filter type; `bpp` = bytes per complete pixel
// (at least 1).  Returns the unfiltered rows packed without filter bytes.
std::vector<uint8_t> png_unfilter(const uint8_t* src, uint32_t rows, size_t stride, int bpp) {
    std::vector<uint8_t> out((size_t)rows * stride);
    std::vector<uint8_t> zero(stride, 0);
    for (uint32_t y = 0; y < rows; ++y) {
        const uint8_t* in = src + (size_t)y * (stride + 1);
        int ft = in[0];
        ++in;
        uint8_t* cur = out.data() + (size_t)y * stride;
        const uint8_t* up = y ? cur - stride : zero.data();
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
            int b = up[i];
            int c = i >= (size_t)bpp ? up[i - bpp] : 0;
            int v = in[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: fail("PNG: invalid filter type");
            }
            cur[i] = (uint8_t)v;
        }
    }
    return out;
}

void decode_png(const uint8_t* d, size_t n, LdrImage& img) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) fail("PNG: bad signature");
    PngHeader hd;
    bool have_ihdr = false, have_trns = false;
    uint8_t pal[256 * 4];
    int pal_n = 0;
    uint32_t key16[3] = {0, 0, 0};
    std::vector<uint8_t> idat;
    size_t pos = 8;
    for (bool end = false; !end;) {
        if (n - pos < 8) fail("PNG: truncated chunk header");
        uint32_t len = be32(d + pos), type = be32(d + pos + 4);
        pos += 8;
        if (len > n - pos || n - pos - len < 4) fail("PNG: truncated chunk");
        const uint8_t* c = d + pos;
        if (type != 0x49484452u && !have_ihdr) fail("PNG: first chunk is not IHDR");
        switch (type) {
            case 0x49484452u:  // IHDR
                if (have_ihdr) fail("PNG: multiple IHDR");
                if (len != 13) fail("PNG: bad IHDR length");
                hd.w = be32(c);
                hd.h = be32(c + 4);
                hd.depth = c[8];
                hd.color = c[9];
                if (!png_depth_ok(hd.color, hd.depth)) fail("PNG: bad colour type / bit depth");
                if (c[10] != 0 || c[11] != 0) fail("PNG: bad compression or filter method");
                hd.interlace = c[12];
                if (hd.interlace > 1) fail("PNG: bad interlace method");
                check_dims(hd.w, hd.h, "PNG");
                hd.channels = hd.color == 0 ? 1 : hd.color == 2 ? 3 : hd.color == 3 ? 1 : hd.color == 4 ? 2 : 4;
                have_ihdr = true;
                break;
            case 0x504c5445u:  // PLTE
                if (len == 0 || len % 3 || len > 768) fail("PNG: bad PLTE length");
                pal_n = (int)(len / 3);
                for (int i = 0; i < pal_n; ++i) put_px(pal + 4 * i, c[3 * i], c[3 * i + 1], c[3 * i + 2], 255);
                break;
            case 0x74524e53u:  // tRNS
                if (!idat.empty()) fail("PNG: tRNS after IDAT");
                if (hd.color == 3) {
                    if (pal_n == 0) fail("PNG: tRNS before PLTE");
                    if (len > (uint32_t)pal_n) fail("PNG: bad tRNS length");
                    for (uint32_t i = 0; i < len; ++i) pal[4 * i + 3] = c[i];
                } else {
                    if (hd.color == 4 || hd.color == 6) fail("PNG: tRNS with alpha");
                    if (len != (uint32_t)hd.channels * 2) fail("PNG: bad tRNS length");
                    for (int k = 0; k < hd.channels; ++k) key16[k] = be16(c + 2 * k);
                    have_trns = true;
                }
                break;
            case 0x49444154u:  // IDAT
                if (hd.color == 3 && pal_n == 0) fail("PNG: no PLTE before IDAT");
                idat.insert(idat.end(), c, c + len);
                break;
            case 0x49454e44u:  // IEND
                end = true;
                break;
            default:
                // Bit 5 of the first type byte clear = critical; CgBI is refused too.
                if (!(type & 0x20000000u) || type == 0x43674249u) fail("PNG: unsupported critical chunk");
                break;
        }
        pos += (size_t)len + 4;  // data + CRC (not verified)
        if (!end && pos >= n) fail("PNG: missing IEND");
    }
    if (idat.empty()) fail("PNG: no image data");

    // Pass geometry: Adam7 or one pass covering the image.
    static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = hd.interlace ? 7 : 1;
    const int bits_pp = hd.channels * hd.depth;
    const int bpp = std::max(1, bits_pp / 8);
    uint32_t pw[7], ph[7];
    size_t stride[7], need = 0;
    for (int p = 0; p < passes; ++p) {
        int x0 = hd.interlace ? ax0[p] : 0, y0 = hd.interlace ? ay0[p] : 0;
        int dx = hd.interlace ? adx[p] : 1, dy = hd.interlace ? ady[p] : 1;
        pw[p] = hd.w > (uint32_t)x0 ? (hd.w - x0 + dx - 1) / dx : 0;
        ph[p] = hd.h > (uint32_t)y0 ? (hd.h - y0 + dy - 1) / dy : 0;
        stride[p] = ((size_t)pw[p] * bits_pp + 7) / 8;
        if (pw[p] && ph[p]) need += (size_t)ph[p] * (stride[p] + 1);
    }
    std::vector<uint8_t> raw(need);
    {
        z_stream zs;
        std::memset(&zs, 0, sizeof zs);
        if (inflateInit(&zs) != Z_OK) fail("PNG: zlib init failed");
        zs.next_in = idat.data();
        zs.avail_in = (uInt)idat.size();
        zs.next_out = raw.data();
        zs.avail_out = (uInt)raw.size();
        int r = inflate(&zs, Z_FINISH);
        size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if (r != Z_STREAM_END && !(got == raw.size() && (r == Z_OK || r == Z_BUF_ERROR)))
            fail(got < raw.size() ? "PNG: not enough pixels" : "PNG: corrupt zlib stream");
        if (got < raw.size()) fail("PNG: not enough pixels");
    }

    img.w = (int)hd.w;
    img.h = (int)hd.h;
    img.rgba.assign((size_t)4 * hd.w * hd.h, 0);
    static const int grey_scale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
    const int gscale = hd.color == 0 ? grey_scale[hd.depth] : 1;
    // The tRNS key at 8 bits is compared after grey scaling, at 16 bits before
    // the high byte is taken.
    uint8_t key8[3];
    for (int k = 0; k < 3; ++k) key8[k] = (uint8_t)((key16[k] & 255u) * (uint32_t)gscale);

    size_t off = 0;
    for (int p = 0; p < passes; ++p) {
        if (!pw[p] || !ph[p]) continue;
        std::vector<uint8_t> rows = png_unfilter(raw.data() + off, ph[p], stride[p], bpp);
        off += (size_t)ph[p] * (stride[p] + 1);
        int x0 = hd.interlace ? ax0[p] : 0, y0 = hd.interlace ? ay0[p] : 0;
        int dx = hd.interlace ? adx[p] : 1, dy = hd.interlace ? ady[p] : 1;
        for (uint32_t y = 0; y < ph[p]; ++y) {
            const uint8_t* row = rows.data() + (size_t)y * stride[p];
            for (uint32_t x = 0; x < pw[p]; ++x) {
                uint32_t s[4];
                for (int k = 0; k < hd.channels; ++k) {
                    size_t bit = ((size_t)x * hd.channels + k) * hd.depth;
                    if (hd.depth == 16) s[k] = be16(row + bit / 8);
                    else if (hd.depth == 8) s[k] = row[bit / 8];
                    else s[k] = (row[bit / 8] >> (8 - hd.depth - (int)(bit % 8))) & ((1u << hd.depth) - 1);
                }
                uint8_t* o = img.rgba.data() + 4 * ((size_t)(y0 + y * dy) * hd.w + (x0 + x * dx));
                const bool d16 = hd.depth == 16;
                switch (hd.color) {
                    case 0: {
                        int g = d16 ? (int)(s[0] >> 8) : (int)(s[0] * gscale);
                        bool clear = have_trns && (d16 ? s[0] == key16[0] : (uint8_t)g == key8[0]);
                        put_px(o, g, g, g, clear ? 0 : 255);
                        break;
                    }
                    case 2: {
                        bool clear = have_trns && (d16 ? (s[0] == key16[0] && s[1] == key16[1] && s[2] == key16[2])
                                                        : (s[0] == key8[0] && s[1] == key8[1] && s[2] == key8[2]));
                        if (d16) put_px(o, s[0] >> 8, s[1] >> 8, s[2] >> 8, clear ? 0 : 255);
                        else put_px(o, s[0], s[1], s[2], clear ? 0 : 255);
                        break;
                    }
                    case 3:
                        if ((int)s[0] >= pal_n) fail("PNG: palette index out of range");
                        std::memcpy(o, pal + 4 * s[0], 4);
                        break;
                    case 4:
                        if (d16) put_px(o, s[0] >> 8, s[0] >> 8, s[0] >> 8, s[1] >> 8);
                        else put_px(o, s[0], s[0], s[0], s[1]);
                        break;
                    default:
                        if (d16) put_px(o, s[0] >> 8, s[1] >> 8, s[2] >> 8, s[3] >> 8);
                        else put_px(o, s[0], s[1], s[2], s[3]);
                        break;
                }
            }
        }
    }
}

// =============================================================== JPEG
// Natural (row-major) index of the k-th coefficient in zig-zag order.
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C) with a 9-bit direct lookup.
struct Huffman {
    bool defined = false;
    uint8_t vals[256] = {};
    int maxcode[17] = {};  // largest code of each length, -1 if none
    int delta[17] = {};    // vals index = code + delta[length]
    uint16_t fast[512] = {};  // (length << 8) | value for codes of <= 9 bits; 0 = longer code

    void build(const uint8_t counts[16], const uint8_t* v, int nv) {
        std::memset(fast, 0, sizeof fast);
        std::memcpy(vals, v, (size_t)nv);
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            int cnt = counts[len - 1];
            delta[len] = k - code;
            for (int i = 0; i < cnt; ++i, ++code, ++k)
                if (len <= 9) {
                    int lo = code << (9 - len), hi = lo + (1 << (9 - len));
                    for (int f = lo; f < hi; ++f) fast[f] = (uint16_t)(len << 8 | vals[k]);
                }
            maxcode[len] = cnt ? code - 1 : -1;
            if (code > (1 << len)) fail("JPEG: bad Huffman table");
            code <<= 1;
        }
        defined = true;
    }
};

// Entropy-coded segment reader: byte stuffing removed, zeros fed once a
// marker is reached.
struct BitReader {
    const uint8_t* p;
    const uint8_t* end;
    uint32_t acc = 0;  // MSB-aligned
    int nbits = 0;
    bool at_marker = false;

    BitReader(const uint8_t* b, const uint8_t* e) : p(b), end(e) {}
    void fill() {
        while (nbits <= 24) {
            uint32_t byte = 0;
            if (!at_marker && p < end) {
                byte = *p;
                if (byte == 0xFF) {
                    const uint8_t* q = p + 1;
                    while (q < end && *q == 0xFF) ++q;
                    if (q < end && *q == 0x00) {
                        p = q + 1;  // stuffed 0xFF data byte
                    } else {
                        at_marker = true;  // p stays on the marker's last 0xFF
                        p = q - 1;
                        byte = 0;
                    }
                } else {
                    ++p;
                }
            }
            acc |= byte << (24 - nbits);
            nbits += 8;
        }
    }
    uint32_t bits(int s) {  // 1 <= s <= 16
        if (nbits < s) fill();
        uint32_t v = acc >> (32 - s);
        acc <<= s;
        nbits -= s;
        return v;
    }
    int bit() { return (int)bits(1); }
    int decode(const Huffman& h) {
        if (nbits < 16) fill();
        uint16_t f = h.fast[acc >> 23];
        if (f) {
            int len = f >> 8;
            acc <<= len;
            nbits -= len;
            return f & 255;
        }
        for (int len = 10; len <= 16; ++len) {
            int code = (int)(acc >> (32 - len));
            if (code <= h.maxcode[len]) {
                acc <<= len;
                nbits -= len;
                return h.vals[code + h.delta[len]];
            }
        }
        fail("JPEG: bad Huffman code");
    }
    // Receive s bits and sign-extend them (T.81 F.2.2.1, EXTEND).
    int receive_extend(int s) {
        if (s == 0) return 0;
        int v = (int)bits(s);
        return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
    }
    // Code of the marker the reader stopped at, 0 when none (or at end of file).
    int marker() const { return at_marker && p + 1 < end ? p[1] : 0; }
    // Position of the next marker (0xFF followed by a non-zero, non-0xFF byte).
    const uint8_t* next_marker() const {
        const uint8_t* q = p;
        while (q + 1 < end && !(q[0] == 0xFF && q[1] != 0x00 && q[1] != 0xFF)) ++q;
        return q;
    }
};

// 1-D LLM inverse DCT on eight integers with 12-bit multipliers.  e0..e3
// are the even-part sums, o0..o3 the odd-part sums; outputs are e_k +- o_k.
struct Llm {
    int e[4], o[4];
};
inline Llm llm_1d(int s0, int s1, int s2, int s3, int s4, int s5, int s6, int s7) {
    constexpr int c0541 = 2217, c_1847 = -7567, c0765 = 3135;
    constexpr int c1175 = 4816, c0298 = 1223, c2053 = 8410, c3072 = 12586, c1501 = 6149;
    constexpr int c_0899 = -3685, c_2562 = -10497, c_1961 = -8034, c_0390 = -1597;
    Llm r;
    int z = (s2 + s6) * c0541;
    int a2 = z + s6 * c_1847, a3 = z + s2 * c0765;
    int b0 = (s0 + s4) * 4096, b1 = (s0 - s4) * 4096;
    r.e[0] = b0 + a3;
    r.e[3] = b0 - a3;
    r.e[1] = b1 + a2;
    r.e[2] = b1 - a2;
    int z71 = s7 + s1, z53 = s5 + s3, z73 = s7 + s3, z51 = s5 + s1;
    int z5 = (z73 + z51) * c1175;
    int w7 = s7 * c0298, w5 = s5 * c2053, w3 = s3 * c3072, w1 = s1 * c1501;
    int y71 = z5 + z71 * c_0899, y53 = z5 + z53 * c_2562;
    int y73 = z73 * c_1961, y51 = z51 * c_0390;
    r.o[0] = w1 + y71 + y51;
    r.o[1] = w3 + y53 + y73;
    r.o[2] = w5 + y53 + y51;
    r.o[3] = w7 + y71 + y73;
    return r;
}

inline uint8_t clamp255(int x) { return (unsigned)x > 255u ? (x < 0 ? 0 : 255) : (uint8_t)x; }

// Dequantised coefficients (natural order) -> 8x8 samples at `out` (stride).
void idct_block(const int16_t* c, uint8_t* out, size_t stride) {
    int tmp[64];
    for (int col = 0; col < 8; ++col) {
        const int16_t* d = c + col;
        if (!(d[8] | d[16] | d[24] | d[32] | d[40] | d[48] | d[56])) {
            for (int r = 0; r < 8; ++r) tmp[8 * r + col] = d[0] * 4;
            continue;
        }
        Llm t = llm_1d(d[0], d[8], d[16], d[24], d[32], d[40], d[48], d[56]);
        for (int k = 0; k < 4; ++k) {
            int e = t.e[k] + 512;
            tmp[8 * k + col] = (e + t.o[k]) >> 10;
            tmp[8 * (7 - k) + col] = (e - t.o[k]) >> 10;
        }
    }
    for (int row = 0; row < 8; ++row) {
        const int* v = tmp + 8 * row;
        Llm t = llm_1d(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
        uint8_t* o = out + (size_t)row * stride;
        for (int k = 0; k < 4; ++k) {
            int e = t.e[k] + 65536 + (128 << 17);
            o[k] = clamp255((e + t.o[k]) >> 17);
            o[7 - k] = clamp255((e - t.o[k]) >> 17);
        }
    }
}

struct JComp {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int x = 0, y = 0;    // samples per line / lines
    int bw = 0, bh = 0;  // blocks per line / column of the MCU-padded plane
    int dc_pred = 0;
    std::vector<uint8_t> plane;  // (8 bw) x (8 bh) samples
    std::vector<int16_t> coef;   // progressive: 64 per block
};

class JpegDecoder {
public:
    JpegDecoder(const uint8_t* d, size_t n) : d_(d), n_(n) {}
    void decode(LdrImage& img);

private:
    const uint8_t* d_;
    size_t n_;
    size_t pos_ = 0;
    uint16_t q_[4][64] = {};  // natural order
    bool qdef_[4] = {};
    Huffman dc_[4], ac_[4];
    JComp comp_[4];
    int nc_ = 0, w_ = 0, h_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
    bool frame_ = false, progressive_ = false, jfif_ = false;
    int app14_ = -1, restart_ = 0, rgb_ids_ = 0;
    int eobrun_ = 0;
    // current scan
    int ns_ = 0, sc_[4] = {}, ss_ = 0, se_ = 0, ah_ = 0, al_ = 0;

    uint8_t byte() {
        if (pos_ >= n_) fail("JPEG: truncated file");
        return d_[pos_++];
    }
    uint32_t word() {
        uint32_t hi = byte();
        return hi << 8 | byte();
    }
    int segment_length() {
        int len = (int)word();
        if (len < 2 || (size_t)(len - 2) > n_ - pos_) fail("JPEG: bad segment length");
        return len - 2;
    }
    void parse_dqt(int len);
    void parse_dht(int len);
    void parse_sof(int len, int marker);
    void parse_sos(int len);
    void decode_scan();
    void decode_block(BitReader& br, JComp& c, int bx, int by);
    void finish_progressive();
    void convert(LdrImage& img);
};

void JpegDecoder::parse_dqt(int len) {
    while (len > 0) {
        int pq_tq = byte();
        int pq = pq_tq >> 4, tq = pq_tq & 15;
        if (pq > 1 || tq > 3) fail("JPEG: bad DQT");
        if (len < 1 + 64 * (pq + 1)) fail("JPEG: bad DQT length");
        for (int k = 0; k < 64; ++k) q_[tq][kZigzag[k]] = (uint16_t)(pq ? word() : byte());
        qdef_[tq] = true;
        len -= 1 + 64 * (pq + 1);
    }
    if (len != 0) fail("JPEG: bad DQT length");
}

void JpegDecoder::parse_dht(int len) {
    while (len > 0) {
        if (len < 17) fail("JPEG: bad DHT length");
        int tc_th = byte();
        int tc = tc_th >> 4, th = tc_th & 15;
        if (tc > 1 || th > 3) fail("JPEG: bad DHT");
        uint8_t counts[16];
        int total = 0;
        for (int i = 0; i < 16; ++i) total += counts[i] = byte();
        if (total > 256 || len < 17 + total) fail("JPEG: bad DHT length");
        uint8_t vals[256];
        for (int i = 0; i < total; ++i) vals[i] = byte();
        (tc ? ac_[th] : dc_[th]).build(counts, vals, total);
        len -= 17 + total;
    }
    if (len != 0) fail("JPEG: bad DHT length");
}

void JpegDecoder::parse_sof(int len, int marker) {
    if (frame_) fail("JPEG: multiple frames");
    if (len < 6) fail("JPEG: bad SOF length");
    if (byte() != 8) fail("JPEG: only 8-bit samples are supported");
    h_ = (int)word();
    w_ = (int)word();
    if (h_ == 0) fail("JPEG: DNL-defined height is not supported");
    check_dims((uint64_t)w_, (uint64_t)h_, "JPEG");
    nc_ = byte();
    if (nc_ != 1 && nc_ != 3 && nc_ != 4) fail("JPEG: bad component count");
    if (len != 6 + 3 * nc_) fail("JPEG: bad SOF length");
    progressive_ = marker == 0xC2;
    static const int rgb_id[3] = {'R', 'G', 'B'};
    for (int i = 0; i < nc_; ++i) {
        JComp& c = comp_[i];
        c.id = byte();
        int hv = byte();
        c.h = hv >> 4;
        c.v = hv & 15;
        c.tq = byte();
        if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) fail("JPEG: bad sampling factors");
        if (c.tq > 3) fail("JPEG: bad quantisation table index");
        if (nc_ == 3 && c.id == rgb_id[i]) ++rgb_ids_;
        hmax_ = std::max(hmax_, c.h);
        vmax_ = std::max(vmax_, c.v);
    }
    for (int i = 0; i < nc_; ++i)
        if (hmax_ % comp_[i].h || vmax_ % comp_[i].v) fail("JPEG: unsupported sampling factors");
    mcux_ = (w_ + 8 * hmax_ - 1) / (8 * hmax_);
    mcuy_ = (h_ + 8 * vmax_ - 1) / (8 * vmax_);
    for (int i = 0; i < nc_; ++i) {
        JComp& c = comp_[i];
        c.x = (w_ * c.h + hmax_ - 1) / hmax_;
        c.y = (h_ * c.v + vmax_ - 1) / vmax_;
        c.bw = mcux_ * c.h;
        c.bh = mcuy_ * c.v;
        c.plane.assign((size_t)64 * c.bw * c.bh, 0);
        if (progressive_) c.coef.assign((size_t)64 * c.bw * c.bh, 0);
    }
    frame_ = true;
}

void JpegDecoder::parse_sos(int len) {
    if (!frame_) fail("JPEG: scan before frame");
    ns_ = byte();
    if (ns_ < 1 || ns_ > nc_ || len != 4 + 2 * ns_) fail("JPEG: bad SOS");
    for (int i = 0; i < ns_; ++i) {
        int id = byte(), tables = byte();
        int k = 0;
        while (k < nc_ && comp_[k].id != id) ++k;
        if (k == nc_) fail("JPEG: scan names an unknown component");
        comp_[k].td = tables >> 4;
        comp_[k].ta = tables & 15;
        if (comp_[k].td > 3 || comp_[k].ta > 3) fail("JPEG: bad SOS table index");
        sc_[i] = k;
    }
    ss_ = byte();
    se_ = byte();
    int a = byte();
    ah_ = a >> 4;
    al_ = a & 15;
    if (progressive_) {
        if (ss_ > 63 || se_ > 63 || ss_ > se_ || ah_ > 13 || al_ > 13) fail("JPEG: bad SOS");
        if (ss_ > 0 && ns_ != 1) fail("JPEG: interleaved AC scan");
    } else {
        if (ss_ != 0 || ah_ != 0 || al_ != 0) fail("JPEG: bad SOS");
        se_ = 63;
    }
}

// One block of the current scan: baseline decode + inverse DCT, or one
// progressive refinement of the stored coefficients (T.81 G.1.2).
void JpegDecoder::decode_block(BitReader& br, JComp& c, int bx, int by) {
    if (!progressive_) {
        const Huffman& hd = dc_[c.td];
        const Huffman& ha = ac_[c.ta];
        if (!hd.defined || !ha.defined) fail("JPEG: missing Huffman table");
        int16_t blk[64] = {};
        const uint16_t* q = q_[c.tq];
        int t = br.decode(hd);
        if (t > 15) fail("JPEG: bad DC category");
        c.dc_pred += br.receive_extend(t);
        blk[0] = (int16_t)(c.dc_pred * (int)q[0]);
        for (int k = 1; k < 64;) {
            int rs = br.decode(ha);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) fail("JPEG: bad AC run");
            int z = kZigzag[k++];
            blk[z] = (int16_t)(br.receive_extend(s) * (int)q[z]);
        }
        idct_block(blk, c.plane.data() + (size_t)8 * by * (8 * c.bw) + 8 * bx, (size_t)8 * c.bw);
        return;
    }
    int16_t* blk = c.coef.data() + (size_t)64 * ((size_t)by * c.bw + bx);
    if (ss_ == 0) {  // DC scan
        if (ah_ == 0) {
            const Huffman& hd = dc_[c.td];
            if (!hd.defined) fail("JPEG: missing Huffman table");
            int t = br.decode(hd);
            if (t > 15) fail("JPEG: bad DC category");
            c.dc_pred += br.receive_extend(t);
            blk[0] = (int16_t)(c.dc_pred * (1 << al_));
        } else if (br.bit()) {
            blk[0] = (int16_t)(blk[0] + (1 << al_));
        }
        return;
    }
    const Huffman& ha = ac_[c.ta];
    if (!ha.defined) fail("JPEG: missing Huffman table");
    if (ah_ == 0) {  // AC first pass
        if (eobrun_ > 0) {
            --eobrun_;
            return;
        }
        for (int k = ss_; k <= se_;) {
            int rs = br.decode(ha);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {
                    eobrun_ = (1 << r) - 1;
                    if (r) eobrun_ += (int)br.bits(r);
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) fail("JPEG: bad AC run");
            blk[kZigzag[k++]] = (int16_t)(br.receive_extend(s) * (1 << al_));
        }
        return;
    }
    // AC refinement: one correction bit per already non-zero coefficient in
    // the band, new +-1 coefficients placed on the r-th zero.
    const int p1 = 1 << al_;
    auto refine = [&](int16_t& v) {
        if (br.bit() && (v & p1) == 0) v = (int16_t)(v >= 0 ? v + p1 : v - p1);
    };
    int k = ss_;
    if (eobrun_ == 0) {
        for (; k <= se_; ++k) {
            int rs = br.decode(ha);
            int r = rs >> 4, s = rs & 15, val = 0;
            if (s) {
                val = br.bit() ? p1 : -p1;
            } else if (r != 15) {
                eobrun_ = 1 << r;
                if (r) eobrun_ += (int)br.bits(r);
                break;
            }
            for (; k <= se_; ++k) {
                int16_t& v = blk[kZigzag[k]];
                if (v != 0) refine(v);
                else if (r-- == 0) break;
            }
            if (val && k <= se_) blk[kZigzag[k]] = (int16_t)val;
        }
    }
    if (eobrun_ > 0) {
        for (; k <= se_; ++k) {
            int16_t& v = blk[kZigzag[k]];
            if (v != 0) refine(v);
        }
        --eobrun_;
    }
}

void JpegDecoder::decode_scan() {
    BitReader br(d_ + pos_, d_ + n_);
    for (int i = 0; i < ns_; ++i) comp_[sc_[i]].dc_pred = 0;
    eobrun_ = 0;
    int todo = restart_ ? restart_ : INT_MAX;
    // Single-component scans cover only the blocks holding samples;
    // interleaved scans cover whole MCUs.
    const bool single = ns_ == 1;
    const int units_x = single ? (comp_[sc_[0]].x + 7) / 8 : mcux_;
    const int units_y = single ? (comp_[sc_[0]].y + 7) / 8 : mcuy_;
    for (int u = 0; u < units_x * units_y; ++u) {
        const int ux = u % units_x, uy = u / units_x;
        if (single) {
            decode_block(br, comp_[sc_[0]], ux, uy);
        } else {
            for (int i = 0; i < ns_; ++i) {
                JComp& c = comp_[sc_[i]];
                for (int yy = 0; yy < c.v; ++yy)
                    for (int xx = 0; xx < c.h; ++xx) decode_block(br, c, ux * c.h + xx, uy * c.v + yy);
            }
        }
        if (--todo <= 0) {
            // End of a restart interval: the entropy data must stop at an
            // RSTn marker here, otherwise the scan ends (its remaining
            // blocks stay as they are).
            br.fill();
            const int mk = br.marker();
            if (mk < 0xD0 || mk > 0xD7) break;
            const uint8_t* m = br.p;
            br = BitReader(m + 2, d_ + n_);
            for (int i = 0; i < ns_; ++i) comp_[sc_[i]].dc_pred = 0;
            eobrun_ = 0;
            todo = restart_ ? restart_ : INT_MAX;
        }
    }
    pos_ = (size_t)((br.at_marker ? br.p : br.next_marker()) - d_);
}

void JpegDecoder::finish_progressive() {
    for (int i = 0; i < nc_; ++i) {
        JComp& c = comp_[i];
        const uint16_t* q = q_[c.tq];
        const int bx_n = (c.x + 7) / 8, by_n = (c.y + 7) / 8;
        for (int by = 0; by < by_n; ++by)
            for (int bx = 0; bx < bx_n; ++bx) {
                int16_t* blk = c.coef.data() + (size_t)64 * ((size_t)by * c.bw + bx);
                int16_t deq[64];
                for (int k = 0; k < 64; ++k) deq[k] = (int16_t)(blk[k] * (int)q[k]);
                idct_block(deq, c.plane.data() + (size_t)8 * by * (8 * c.bw) + 8 * bx, (size_t)8 * c.bw);
            }
    }
}

// Upsample every component to full resolution and convert to RGBA.
void JpegDecoder::convert(LdrImage& img) {
    img.w = w_;
    img.h = h_;
    img.rgba.assign((size_t)4 * w_ * h_, 0);
    const bool is_rgb = nc_ == 3 && (rgb_ids_ == 3 || (app14_ == 0 && !jfif_));
    std::vector<uint8_t> line[4];
    const uint8_t* row[4];
    for (int k = 0; k < nc_; ++k) line[k].assign((size_t)w_ + 8 * hmax_, 0);
    for (int j = 0; j < h_; ++j) {
        for (int k = 0; k < nc_; ++k) {
            const JComp& c = comp_[k];
            const int hs = hmax_ / c.h, vs = vmax_ / c.v;
            const int wl = (w_ + hs - 1) / hs;  // samples of this component per line
            const size_t pitch = (size_t)8 * c.bw;
            const int ny = std::min(j / vs, c.y - 1);
            int fy = ny;
            if (vs == 2) fy = (j & 1) ? std::min(j / 2 + 1, c.y - 1) : std::max(j / 2 - 1, 0);
            const uint8_t* nr = c.plane.data() + (size_t)ny * pitch;
            const uint8_t* fr = c.plane.data() + (size_t)fy * pitch;
            uint8_t* o = line[k].data();
            if (hs == 1 && vs == 1) {
                row[k] = nr;
                continue;
            }
            row[k] = o;
            if (hs == 1 && vs == 2) {
                for (int i = 0; i < wl; ++i) o[i] = (uint8_t)((3 * nr[i] + fr[i] + 2) >> 2);
            } else if (hs == 2 && vs == 1) {
                if (wl == 1) {
                    o[0] = o[1] = nr[0];
                    continue;
                }
                o[0] = nr[0];
                o[1] = (uint8_t)((3 * nr[0] + nr[1] + 2) >> 2);
                for (int i = 1; i < wl - 1; ++i) {
                    int m = 3 * nr[i] + 2;
                    o[2 * i] = (uint8_t)((m + nr[i - 1]) >> 2);
                    o[2 * i + 1] = (uint8_t)((m + nr[i + 1]) >> 2);
                }
                // The last pair: the left sample weights the previous input by 3/4.
                o[2 * wl - 2] = (uint8_t)((3 * nr[wl - 2] + nr[wl - 1] + 2) >> 2);
                o[2 * wl - 1] = nr[wl - 1];
            } else if (hs == 2 && vs == 2) {
                int t1 = 3 * nr[0] + fr[0];
                if (wl == 1) {
                    o[0] = o[1] = (uint8_t)((t1 + 2) >> 2);
                    continue;
                }
                o[0] = (uint8_t)((t1 + 2) >> 2);
                for (int i = 1; i < wl; ++i) {
                    int t0 = t1;
                    t1 = 3 * nr[i] + fr[i];
                    o[2 * i - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
                    o[2 * i] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
                }
                o[2 * wl - 1] = (uint8_t)((t1 + 2) >> 2);
            } else {
                for (int i = 0; i < wl; ++i)
                    for (int s = 0; s < hs; ++s) o[i * hs + s] = nr[i];
            }
        }
        uint8_t* out = img.rgba.data() + (size_t)4 * w_ * j;
        auto ycc = [](int y, int cb, int cr, uint8_t* o) {
            int yf = (y << 20) + (1 << 19);
            cr -= 128;
            cb -= 128;
            int r = (yf + cr * (5743 << 8)) >> 20;
            int g = (yf + cr * -(2925 << 8) + (int)((unsigned)(cb * -(1410 << 8)) & 0xffff0000u)) >> 20;
            int b = (yf + cb * (7258 << 8)) >> 20;
            put_px(o, clamp255(r), clamp255(g), clamp255(b), 255);
        };
        auto blend = [](int x, int m) {  // x * m / 255, rounded
            unsigned t = (unsigned)(x * m + 128);
            return (int)((t + (t >> 8)) >> 8);
        };
        for (int i = 0; i < w_; ++i, out += 4) {
            if (nc_ == 1) {
                put_px(out, row[0][i], row[0][i], row[0][i], 255);
            } else if (nc_ == 3) {
                if (is_rgb) put_px(out, row[0][i], row[1][i], row[2][i], 255);
                else ycc(row[0][i], row[1][i], row[2][i], out);
            } else if (app14_ == 0) {  // CMYK
                int m = row[3][i];
                put_px(out, blend(row[0][i], m), blend(row[1][i], m), blend(row[2][i], m), 255);
            } else if (app14_ == 2) {  // YCCK
                ycc(row[0][i], row[1][i], row[2][i], out);
                int m = row[3][i];
                put_px(out, blend(255 - out[0], m), blend(255 - out[1], m), blend(255 - out[2], m), 255);
            } else {  // YCbCr with a fourth channel that is ignored
                ycc(row[0][i], row[1][i], row[2][i], out);
            }
        }
    }
}

void JpegDecoder::decode(LdrImage& img) {
    if (n_ < 4 || d_[0] != 0xFF || d_[1] != 0xD8) fail("JPEG: missing SOI");
    pos_ = 2;
    int scans = 0;
    for (;;) {
        // Next marker: skip anything up to 0xFF, then fill bytes.
        while (pos_ < n_ && d_[pos_] != 0xFF) ++pos_;
        while (pos_ < n_ && d_[pos_] == 0xFF) ++pos_;
        if (pos_ >= n_) {
            if (scans == 0) fail("JPEG: no scan before end of file");
            break;
        }
        int m = d_[pos_++];
        if (m == 0xD9) break;                      // EOI
        if (m >= 0xD0 && m <= 0xD7) continue;      // stray RSTn
        if (m == 0x01 || m == 0xD8) fail("JPEG: unexpected marker");
        int len = segment_length();
        size_t seg_end = pos_ + (size_t)len;
        if (m == 0xDB) {
            parse_dqt(len);
        } else if (m == 0xC4) {
            parse_dht(len);
        } else if (m == 0xDD) {
            if (len != 2) fail("JPEG: bad DRI length");
            restart_ = (int)word();
        } else if (m == 0xC0 || m == 0xC1 || m == 0xC2) {
            parse_sof(len, m);
        } else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            fail("JPEG: unsupported coding process (lossless, hierarchical or arithmetic)");
        } else if (m == 0xDC) {
            fail("JPEG: DNL marker is not supported");
        } else if (m == 0xDA) {
            parse_sos(len);
            for (int i = 0; i < nc_; ++i)
                if (!qdef_[comp_[i].tq]) fail("JPEG: missing quantisation table");
            decode_scan();
            ++scans;
            continue;  // pos_ is at the marker after the scan
        } else if (m == 0xE0 && len >= 5 && std::memcmp(d_ + pos_, "JFIF\0", 5) == 0) {
            jfif_ = true;
        } else if (m == 0xEE && len >= 12 && std::memcmp(d_ + pos_, "Adobe\0", 6) == 0) {
            app14_ = d_[pos_ + 11];
        }
        pos_ = seg_end;
    }
    if (!frame_) fail("JPEG: no frame header");
    if (progressive_) finish_progressive();
    convert(img);
}

// =============================================================== PNM
void decode_pnm(const uint8_t* d, size_t n, LdrImage& img) {
    const int ch = d[1] == '5' ? 1 : 3;
    size_t pos = 2;
    auto number = [&]() -> uint64_t {
        for (;;) {  // whitespace and comments
            while (pos < n && (d[pos] == ' ' || (d[pos] >= 9 && d[pos] <= 13))) ++pos;
            if (pos < n && d[pos] == '#') {
                while (pos < n && d[pos] != '\n' && d[pos] != '\r') ++pos;
                continue;
            }
            break;
        }
        if (pos >= n || d[pos] < '0' || d[pos] > '9') fail("PNM: bad header");
        uint64_t v = 0;
        while (pos < n && d[pos] >= '0' && d[pos] <= '9') {
            v = v * 10 + (uint64_t)(d[pos++] - '0');
            if (v > (1ull << 32)) fail("PNM: bad header");
        }
        return v;
    };
    uint64_t w = number(), h = number(), maxv = number();
    check_dims(w, h, "PNM");
    if (maxv == 0 || maxv > 65535) fail("PNM: bad maxval");
    ++pos;  // the single whitespace byte after maxval
    const int bytes = maxv > 255 ? 2 : 1;
    const size_t need = (size_t)w * h * ch * bytes;
    if (pos > n || n - pos < need) fail("PNM: truncated data");
    img.w = (int)w;
    img.h = (int)h;
    img.rgba.assign((size_t)4 * w * h, 0);
    const uint8_t* s = d + pos;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        int v[3];
        // 16-bit samples: the reference's loader reads them in little-endian
        // host order and keeps the upper half, i.e. the second file byte.
        for (int k = 0; k < ch; ++k) v[k] = s[((size_t)i * ch + k) * bytes + (bytes - 1)];
        if (ch == 1) v[1] = v[2] = v[0];
        put_px(img.rgba.data() + 4 * i, v[0], v[1], v[2], 255);
    }
}

}  // namespace

void decode_ldr_image(const uint8_t* data, size_t size, LdrImage& out) {
    if (size >= 8 && data[0] == 0x89 && data[1] == 'P' && data[2] == 'N' && data[3] == 'G') {
        decode_png(data, size, out);
    } else if (size >= 3 && data[0] == 0xFF && data[1] == 0xD8) {
        JpegDecoder(data, size).decode(out);
    } else if (size >= 3 && data[0] == 'P' && (data[1] == '5' || data[1] == '6')) {
        decode_pnm(data, size, out);
    } else {
        fail("unsupported image format (PNG, JPEG and binary PNM are read)");
    }
}

void write_png_rgba8(const std::string& path, int w, int h, const uint8_t* rgba) {
    if (w <= 0 || h <= 0) fail("PNG: empty image");
    const size_t stride = (size_t)4 * w;
    std::vector<uint8_t> raw((stride + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw[(stride + 1) * y] = 0;  // no row filter
        std::memcpy(&raw[(stride + 1) * y + 1], rgba + stride * y, stride);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) fail("PNG: zlib compression failed");
    std::vector<uint8_t> file = {137, 80, 78, 71, 13, 10, 26, 10};
    auto chunk = [&](const char* tag, const uint8_t* data, size_t len) {
        const uint8_t hdr[8] = {(uint8_t)(len >> 24), (uint8_t)(len >> 16), (uint8_t)(len >> 8), (uint8_t)len,
                                (uint8_t)tag[0], (uint8_t)tag[1], (uint8_t)tag[2], (uint8_t)tag[3]};
        file.insert(file.end(), hdr, hdr + 8);
        file.insert(file.end(), data, data + len);
        uLong crc = crc32(0L, hdr + 4, 4);
        crc = crc32(crc, data, (uInt)len);
        const uint8_t c[4] = {(uint8_t)(crc >> 24), (uint8_t)(crc >> 16), (uint8_t)(crc >> 8), (uint8_t)crc};
        file.insert(file.end(), c, c + 4);
    };
    const uint8_t ihdr[13] = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w,
                              (uint8_t)(h >> 24), (uint8_t)(h >> 16), (uint8_t)(h >> 8), (uint8_t)h,
                              8, 6, 0, 0, 0};
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), zlen);
    chunk("IEND", nullptr, 0);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) fail("cannot write '" + path + "'");
    bool ok = std::fwrite(file.data(), 1, file.size(), f) == file.size();
    ok = std::fclose(f) == 0 && ok;
    if (!ok) fail("cannot write '" + path + "'");
}

void write_pfm_rgb(const std::string& path, int w, int h, const float* rgba) {
    if (w <= 0 || h <= 0) fail("PFM: empty image");
    std::string hdr = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    std::vector<float> rgb((size_t)3 * w * h);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int k = 0; k < 3; ++k) rgb[3 * i + k] = rgba[4 * i + k];
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) fail("cannot write '" + path + "'");
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
              std::fwrite(rgb.data(), sizeof(float), rgb.size(), f) == rgb.size();  // little-endian host
    ok = std::fclose(f) == 0 && ok;
    if (!ok) fail("cannot write '" + path + "'");
}

void load_ldr_image(const std::string& path, LdrImage& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("cannot open image '" + path + "'");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    try {
        decode_ldr_image(bytes.data(), bytes.size(), out);
    } catch (const std::exception& e) {
        fail("image '" + path + "': " + e.what());
    }
}

}  // namespace rt
