#include "image_out.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace srt {

uint8_t to_ldr(float x) {
    const float c = std::min(std::max(x, 0.0f), 1.0f) * 255.0f;
    return (uint8_t)c;
}

// float -> IEEE half, round to nearest even (overflow to infinity, NaN kept)
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);               // rounds to >= 65520: infinity
    if (ax < 0x38800000u) {                                                   // below the smallest normal half
        if (ax < 0x33000000u) return (uint16_t)sign;                          // below half the smallest subnormal
        const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - e;                                      // 14..24
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((ax - 0x38000000u) >> 13);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

// OpenEXR 2.0 scanline file, no compression, HALF channels (the reference
// writes HALF, texture.cu:96-117); channel names in the file's required
// alphabetical order A, B, G, R; rows top to bottom.
void write_exr(std::ostream& f, int W, int H, const PixelFn& px) {
    std::string hdr;
    auto put32 = [&](std::string& s, int32_t v) { s.append((const char*)&v, 4); };
    auto attr = [&](const char* name, const char* type, const std::string& value) {
        hdr += name; hdr += '\0';
        hdr += type; hdr += '\0';
        put32(hdr, (int32_t)value.size());
        hdr += value;
    };
    std::string ch;
    for (const char* name : {"A", "B", "G", "R"}) {
        ch += name; ch += '\0';
        put32(ch, 1);                        // HALF
        ch += std::string(4, '\0');           // pLinear + reserved
        put32(ch, 1); put32(ch, 1);          // x / y sampling
    }
    ch += '\0';
    std::string box;
    put32(box, 0); put32(box, 0); put32(box, W - 1); put32(box, H - 1);
    const float one = 1.0f, zero = 0.0f;
    attr("channels", "chlist", ch);
    attr("compression", "compression", std::string(1, '\0'));
    attr("dataWindow", "box2i", box);
    attr("displayWindow", "box2i", box);
    attr("lineOrder", "lineOrder", std::string(1, '\0'));
    attr("pixelAspectRatio", "float", std::string((const char*)&one, 4));
    std::string center((const char*)&zero, 4);
    center.append((const char*)&zero, 4);
    attr("screenWindowCenter", "v2f", center);
    attr("screenWindowWidth", "float", std::string((const char*)&one, 4));
    hdr += '\0';
    const uint32_t magic = 20000630u, version = 2u;
    f.write((const char*)&magic, 4);
    f.write((const char*)&version, 4);
    f.write(hdr.data(), (std::streamsize)hdr.size());
    const uint64_t line_bytes = 8ull * W;    // 4 channels x 2 bytes
    const uint64_t start = 8 + hdr.size() + 8ull * H;
    for (int y = 0; y < H; y++) {
        const uint64_t off = start + (uint64_t)y * (8 + line_bytes);
        f.write((const char*)&off, 8);
    }
    std::vector<uint16_t> line(4 * (size_t)W);
    static const int chan[4] = {3, 2, 1, 0};    // A, B, G, R from RGBA
    for (int y = 0; y < H; y++) {
        for (int k = 0; k < 4; k++)
            for (int x = 0; x < W; x++) line[(size_t)k * W + x] = to_half(px(y, x)[chan[k]]);
        const int32_t yy = y, size = (int32_t)line_bytes;
        f.write((const char*)&yy, 4);
        f.write((const char*)&size, 4);
        f.write((const char*)line.data(), (std::streamsize)line_bytes);
    }
}


// ---------------------------------------------------------------------------
// Radiance RGBE (.hdr): the "#?RADIANCE" header, then the rows top to bottom.
// A pixel is a shared exponent e and three 8-bit mantissas: for v = max(r, g,
// b) = m 2^e with m in [0.5, 1), the mantissas are c m 256 / v truncated and the
// exponent byte is e + 128 (Ward, "Real Pixels", Graphics Gems II, 1991).
// Rows 8..32767 wide use the per-channel run-length layout (marker 2 2 W), here
// with literal runs only; other widths are written flat.
// ---------------------------------------------------------------------------
static void rgbe(const float* p, uint8_t out[4]) {
    const float r = std::max(p[0], 0.0f), g = std::max(p[1], 0.0f), b = std::max(p[2], 0.0f);
    const float v = std::max(r, std::max(g, b));
    if (!(v >= 1e-32f)) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e = 0;
    const float m = std::frexp(v, &e);
    const float scale = m * 256.0f / v;
    out[0] = (uint8_t)std::min(r * scale, 255.0f);
    out[1] = (uint8_t)std::min(g * scale, 255.0f);
    out[2] = (uint8_t)std::min(b * scale, 255.0f);
    out[3] = (uint8_t)(e + 128);
}

void write_hdr(std::ostream& f, int W, int H, const PixelFn& px) {
    f << "#?RADIANCE\n# srtracer\nFORMAT=32-bit_rle_rgbe\n\n-Y " << H << " +X " << W << "\n";
    std::vector<uint8_t> row(4 * (size_t)W);
    std::vector<uint8_t> out;
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) rgbe(px(y, x), &row[4 * (size_t)x]);
        out.clear();
        if (W >= 8 && W <= 0x7fff) {
            out.push_back(2); out.push_back(2); out.push_back((uint8_t)(W >> 8)); out.push_back((uint8_t)(W & 255));
            for (int c = 0; c < 4; c++)
                for (int x0 = 0; x0 < W; x0 += 128) {
                    const int n = std::min(128, W - x0);
                    out.push_back((uint8_t)n);
                    for (int x = x0; x < x0 + n; x++) out.push_back(row[4 * (size_t)x + c]);
                }
        } else {
            out = row;
        }
        f.write((const char*)out.data(), (std::streamsize)out.size());
    }
}

// ---------------------------------------------------------------------------
// PNG: 8-bit RGBA, filter 0 on every row, one IDAT holding a zlib stream of
// stored (uncompressed) deflate blocks (RFC 2083, RFC 1950, RFC 1951 3.2.4).
// ---------------------------------------------------------------------------
static uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
    return crc;
}

static void put_be32(std::string& s, uint32_t v) {
    const char b[4] = {(char)(v >> 24), (char)(v >> 16), (char)(v >> 8), (char)v};
    s.append(b, 4);
}

static void png_chunk(std::ostream& f, const char* type, const std::string& data) {
    std::string c;
    put_be32(c, (uint32_t)data.size());
    c.append(type, 4);
    c += data;
    const uint32_t crc = crc32_update(0xffffffffu, (const uint8_t*)c.data() + 4, c.size() - 4) ^ 0xffffffffu;
    put_be32(c, crc);
    f.write(c.data(), (std::streamsize)c.size());
}

void write_png(std::ostream& f, int W, int H, const PixelFn& px) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    f.write((const char*)sig, 8);
    std::string ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr += (char)8;    // bit depth
    ihdr += (char)6;    // RGBA
    ihdr += std::string(3, '\0');    // deflate, adaptive filtering, no interlace
    png_chunk(f, "IHDR", ihdr);
    // filtered image data: a filter-type byte (0) before every row
    std::string raw;
    raw.reserve((size_t)H * (1 + 4 * (size_t)W));
    for (int y = 0; y < H; y++) {
        raw += '\0';
        for (int x = 0; x < W; x++) {
            const float* p = px(y, x);
            for (int c = 0; c < 4; c++) raw += (char)to_ldr(p[c]);
        }
    }
    std::string z;
    z += (char)0x78;
    z += (char)0x01;    // deflate, 32 KiB window, no preset dictionary; (0x78 << 8 | 0x01) % 31 == 0
    const size_t n = raw.size();
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, n - pos);
        const bool last = pos + len == n;
        z += (char)(last ? 1 : 0);    // BFINAL, BTYPE 00 (stored)
        z += (char)(len & 0xff);
        z += (char)(len >> 8);
        z += (char)(~len & 0xff);
        z += (char)((~len >> 8) & 0xff);
        z.append(raw, pos, len);
        pos += len;
    } while (pos < n);
    uint32_t a = 1, b = 0;    // Adler-32
    for (size_t i = 0; i < n; i++) {
        a = (a + (uint8_t)raw[i]) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    png_chunk(f, "IDAT", z);
    png_chunk(f, "IEND", std::string());
}

// ---------------------------------------------------------------------------
// BMP: 32 bits per pixel with a BITMAPV4HEADER and BI_BITFIELDS masks for
// B, G, R, A bytes, rows bottom to top (the layout the reference's writer
// uses for 4 channels).
// ---------------------------------------------------------------------------
void write_bmp(std::ostream& f, int W, int H, const PixelFn& px) {
    std::string h;
    auto le16 = [&](uint32_t v) { h += (char)(v & 0xff); h += (char)((v >> 8) & 0xff); };
    auto le32 = [&](uint32_t v) { for (int k = 0; k < 4; k++) h += (char)((v >> (8 * k)) & 0xff); };
    const uint32_t data = 4u * (uint32_t)W * (uint32_t)H, off = 14u + 108u;
    h += 'B'; h += 'M';
    le32(off + data); le16(0); le16(0); le32(off);
    le32(108); le32((uint32_t)W); le32((uint32_t)H); le16(1); le16(32);
    le32(3);             // BI_BITFIELDS
    le32(data); le32(2835); le32(2835); le32(0); le32(0);
    le32(0x00ff0000u); le32(0x0000ff00u); le32(0x000000ffu); le32(0xff000000u);
    le32(0x73524742u);   // 'sRGB' colour space
    for (int k = 0; k < 12; k++) le32(0);    // endpoints (unused for sRGB), gamma
    f.write(h.data(), (std::streamsize)h.size());
    std::vector<uint8_t> row(4 * (size_t)W);
    for (int y = H - 1; y >= 0; y--) {
        for (int x = 0; x < W; x++) {
            const float* p = px(y, x);
            row[4 * (size_t)x + 0] = to_ldr(p[2]);
            row[4 * (size_t)x + 1] = to_ldr(p[1]);
            row[4 * (size_t)x + 2] = to_ldr(p[0]);
            row[4 * (size_t)x + 3] = to_ldr(p[3]);
        }
        f.write((const char*)row.data(), (std::streamsize)row.size());
    }
}

// ---------------------------------------------------------------------------
// JPEG: baseline sequential DCT (ITU-T T.81), JFIF, YCbCr 4:4:4, quality 100
// (every quantiser 1, as the reference asks of its writer).  One quantisation
// table and one DC / AC Huffman table pair serve the three components; the
// Huffman tables are fixed-length canonical codes (4 bits for the 12 DC
// categories, 8 bits for the 162 AC run/size symbols), valid for any data.
// ---------------------------------------------------------------------------
namespace {

struct BitWriter {
    std::string out;
    uint32_t acc = 0;
    int nbits = 0;
    void put(uint32_t code, int len) {
        for (int i = len - 1; i >= 0; i--) {
            acc = (acc << 1) | ((code >> i) & 1u);
            if (++nbits == 8) {
                out += (char)acc;
                if (acc == 0xffu) out += '\0';    // byte stuffing
                acc = 0;
                nbits = 0;
            }
        }
    }
    void flush() {
        while (nbits != 0) put(1u, 1);    // pad with 1 bits
    }
};

int category(int v) {
    int a = v < 0 ? -v : v, n = 0;
    while (a) {
        n++;
        a >>= 1;
    }
    return n;
}

// the value bits that follow a category: v itself, or v - 1 for v < 0, in n bits
uint32_t value_bits(int v, int n) { return (uint32_t)(v < 0 ? v + (1 << n) - 1 : v) & ((1u << n) - 1u); }

}  // namespace

void write_jpg(std::ostream& f, int W, int H, const PixelFn& px) {
    if (W > 65535 || H > 65535) return;
    // zigzag order: entry k of the scan is coefficient zz[k] (row-major index)
    int zz[64];
    {
        int k = 0;
        for (int s = 0; s < 15; s++) {
            if (s % 2 == 0) {    // up and to the right: row decreases
                for (int r = std::min(s, 7); r >= 0 && s - r <= 7; r--) zz[k++] = r * 8 + (s - r);
            } else {
                for (int c = std::min(s, 7); c >= 0 && s - c <= 7; c--) zz[k++] = (s - c) * 8 + c;
            }
        }
    }
    // AC symbols: EOB, ZRL, then run/size pairs; code of symbol i = i (8 bits)
    std::vector<uint8_t> ac_syms = {0x00, 0xf0};
    for (int r = 0; r < 16; r++)
        for (int sz = 1; sz <= 10; sz++) ac_syms.push_back((uint8_t)((r << 4) | sz));
    int ac_code[256];
    for (int i = 0; i < 256; i++) ac_code[i] = -1;
    for (size_t i = 0; i < ac_syms.size(); i++) ac_code[ac_syms[i]] = (int)i;

    std::string hdr;
    auto b8 = [&](int v) { hdr += (char)(v & 0xff); };
    auto b16 = [&](int v) { b8(v >> 8); b8(v); };
    b16(0xffd8);
    b16(0xffe0); b16(16); hdr += std::string("JFIF", 4); b8(0); b8(1); b8(1); b8(0); b16(1); b16(1); b8(0); b8(0);
    b16(0xffdb); b16(67); b8(0x00);
    for (int i = 0; i < 64; i++) b8(1);
    b16(0xffc0); b16(17); b8(8); b16(H); b16(W); b8(3);
    for (int c = 1; c <= 3; c++) { b8(c); b8(0x11); b8(0); }
    b16(0xffc4); b16(2 + 17 + 12 + 17 + 162);
    b8(0x00);
    for (int l = 1; l <= 16; l++) b8(l == 4 ? 12 : 0);
    for (int v = 0; v < 12; v++) b8(v);
    b8(0x10);
    for (int l = 1; l <= 16; l++) b8(l == 8 ? 162 : 0);
    for (uint8_t v : ac_syms) b8(v);
    b16(0xffda); b16(12); b8(3);
    for (int c = 1; c <= 3; c++) { b8(c); b8(0x00); }
    b8(0); b8(63); b8(0);
    f.write(hdr.data(), (std::streamsize)hdr.size());

    // DCT basis: F(u, v) = sum_x sum_y cb[u][x] cb[v][y] f(x, y), cb[u][x] = C(u) / 2 cos((2x + 1) u pi / 16)
    double cb[8][8];
    for (int u = 0; u < 8; u++)
        for (int x = 0; x < 8; x++)
            cb[u][x] = (u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0);
    // YCbCr planes of the 8-bit image (JFIF conversion)
    std::vector<float> plane[3];
    for (auto& p : plane) p.resize((size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* p = px(y, x);
            const float r = to_ldr(p[0]), g = to_ldr(p[1]), b = to_ldr(p[2]);
            const size_t i = (size_t)y * W + x;
            plane[0][i] = 0.299f * r + 0.587f * g + 0.114f * b;
            plane[1][i] = -0.168736f * r - 0.331264f * g + 0.5f * b + 128.0f;
            plane[2][i] = 0.5f * r - 0.418688f * g - 0.081312f * b + 128.0f;
        }
    BitWriter bw;
    int pred[3] = {0, 0, 0};
    double blk[64], tmp[64];
    for (int by = 0; by < H; by += 8)
        for (int bx = 0; bx < W; bx += 8)
            for (int c = 0; c < 3; c++) {
                for (int y = 0; y < 8; y++)
                    for (int x = 0; x < 8; x++) {    // edge pixels repeat past the image
                        const int sy = std::min(by + y, H - 1), sx = std::min(bx + x, W - 1);
                        blk[y * 8 + x] = plane[c][(size_t)sy * W + sx] - 128.0;
                    }
                for (int y = 0; y < 8; y++)
                    for (int u = 0; u < 8; u++) {
                        double s = 0.0;
                        for (int x = 0; x < 8; x++) s += cb[u][x] * blk[y * 8 + x];
                        tmp[y * 8 + u] = s;
                    }
                int q[64];
                for (int v = 0; v < 8; v++)
                    for (int u = 0; u < 8; u++) {
                        double s = 0.0;
                        for (int y = 0; y < 8; y++) s += cb[v][y] * tmp[y * 8 + u];
                        // DC within [-1024, 1023] (DC differences <= 11 bits), AC within +-1023 (<= 10 bits)
                        const int lo = (u == 0 && v == 0) ? -1024 : -1023;
                        q[v * 8 + u] = std::max(lo, std::min(1023, (int)std::lround(s)));
                    }
                const int diff = q[0] - pred[c];
                pred[c] = q[0];
                const int dn = category(diff);
                bw.put((uint32_t)dn, 4);
                if (dn) bw.put(value_bits(diff, dn), dn);
                int run = 0;
                for (int k = 1; k < 64; k++) {
                    const int v = q[zz[k]];
                    if (v == 0) {
                        run++;
                        continue;
                    }
                    while (run > 15) {
                        bw.put((uint32_t)ac_code[0xf0], 8);
                        run -= 16;
                    }
                    const int n = category(v);
                    bw.put((uint32_t)ac_code[(run << 4) | n], 8);
                    bw.put(value_bits(v, n), n);
                    run = 0;
                }
                if (run > 0) bw.put((uint32_t)ac_code[0x00], 8);
            }
    bw.flush();
    f.write(bw.out.data(), (std::streamsize)bw.out.size());
    const char eoi[2] = {(char)0xff, (char)0xd9};
    f.write(eoi, 2);
}

}  // namespace srt
