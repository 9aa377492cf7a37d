// fits_io.hpp -- a small FITS (NOST 100-2.0) image reader/writer.
//
// Enough of the standard for the gridding path's files: primary and IMAGE
// extension HDUs, BITPIX 8/16/32/64/-32/-64 with BSCALE/BZERO, big-endian
// data in 2880-byte blocks.  It replaces the cfitsio calls of
// HCGrid/gmap.cpp (fits_open_file, fits_read_img, fits_hdr2str,
// fits_create_img, fits_write_record, fits_write_img) without a library
// dependency.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace hcgio {

struct Hdu {
  std::vector<std::string> cards;  // 80-character records, END excluded
  int bitpix = 0;
  std::vector<int64_t> naxes;
  double bscale = 1.0, bzero = 0.0;
  uint64_t data_offset = 0;  // bytes from the start of the file
  uint64_t nelem = 0;
};

// Keyword lookup over a card list (first occurrence).
bool card_string(const std::vector<std::string>& cards, const std::string& key, std::string& out);
bool card_double(const std::vector<std::string>& cards, const std::string& key, double& out);
bool card_int(const std::vector<std::string>& cards, const std::string& key, int64_t& out);

class FitsFile {
 public:
  bool open(const std::string& path, std::string& err);
  size_t num_hdus() const { return hdus_.size(); }
  const Hdu& hdu(size_t i) const { return hdus_[i]; }
  // first is 1-based like fits_read_img; converts to double with BSCALE/BZERO
  bool read(size_t hdu, uint64_t first, uint64_t count, double* out, std::string& err) const;
  // all cards of an HDU plus the END card, concatenated (what fits_hdr2str returns)
  std::string header_string(size_t hdu) const;

 private:
  std::string path_;
  std::vector<Hdu> hdus_;
  std::vector<unsigned char> bytes_;
};

// One HDU to write: its cards (mandatory ones are generated), BITPIX, axes, data.
struct HduOut {
  int bitpix = -64;
  std::vector<int64_t> naxes;
  std::vector<std::string> extra_cards;  // appended after the mandatory cards
  const double* data = nullptr;          // nelem values (fewer -> zero fill)
  uint64_t ndata = 0;
};

// Writes a primary HDU and optional IMAGE extensions.  A leading '!' in the
// path means "overwrite" (cfitsio's convention, HCGrid/HCGrid.cpp:112-130).
bool write_file(const std::string& path, const std::vector<HduOut>& hdus, std::string& err);

// Pads/cuts a record to 80 characters.
std::string card80(const std::string& s);

}  // namespace hcgio
