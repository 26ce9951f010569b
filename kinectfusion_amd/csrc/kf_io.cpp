// Host-side image codec helper of the reconstruction app: reversal of the
// PNG row filters (ISO/IEC 15948 section 9: None, Sub, Up, Average, Paeth)
// after the zlib stream has been inflated.  The per-byte recurrences of Sub,
// Average and Paeth run left to right along a row and from the row above, so
// they are a sequential loop; the rest of the PNG container is handled by
// kinectfusion_amd/app/png.py.
#include <cstdint>
#include <cstdlib>

extern "C" {

// in:  height rows of (1 filter-type byte + row_bytes filtered bytes);
// out: height * row_bytes reconstructed bytes; bpp: bytes per complete pixel
// (at least 1).  Returns 0, or -1 on an unknown filter type.
int kf_png_unfilter(const std::uint8_t* in, std::uint8_t* out, int height, int row_bytes, int bpp) {
  const std::uint8_t* prev = nullptr;
  for (int y = 0; y < height; ++y) {
    const std::uint8_t* src = in + static_cast<std::size_t>(y) * (row_bytes + 1);
    const int type = src[0];
    ++src;
    std::uint8_t* dst = out + static_cast<std::size_t>(y) * row_bytes;
    for (int i = 0; i < row_bytes; ++i) {
      const int a = i >= bpp ? dst[i - bpp] : 0;
      const int b = prev != nullptr ? prev[i] : 0;
      const int c = (prev != nullptr && i >= bpp) ? prev[i - bpp] : 0;
      int pred;
      switch (type) {
        case 0:
          pred = 0;
          break;
        case 1:
          pred = a;
          break;
        case 2:
          pred = b;
          break;
        case 3:
          pred = (a + b) >> 1;
          break;
        case 4: {
          const int p = a + b - c;
          const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
          pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
          break;
        }
        default:
          return -1;
      }
      dst[i] = static_cast<std::uint8_t>(src[i] + pred);
    }
    prev = dst;
  }
  return 0;
}

}  // extern "C"
