// fits_io.cpp -- see fits_io.hpp.
#include "fits_io.hpp"

#include <stdio.h>
#include <string.h>

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace hcgio {

namespace {
constexpr size_t kBlock = 2880;
constexpr size_t kCard = 80;

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(' ');
  return s.substr(a, b - a + 1);
}

// Value field of a card: text after "= " up to a '/' comment outside quotes.
bool card_value(const std::string& card, std::string& v) {
  if (card.size() < 10 || card[8] != '=' || card[9] != ' ') return false;
  std::string rest = card.substr(10);
  size_t i = rest.find_first_not_of(' ');
  if (i == std::string::npos) return false;
  if (rest[i] == '\'') {
    std::string out;
    for (size_t j = i + 1; j < rest.size(); ++j) {
      if (rest[j] == '\'') {
        if (j + 1 < rest.size() && rest[j + 1] == '\'') {
          out.push_back('\'');
          ++j;
          continue;
        }
        v = trim(out);
        return true;
      }
      out.push_back(rest[j]);
    }
    return false;
  }
  size_t slash = rest.find('/', i);
  v = trim(rest.substr(i, slash == std::string::npos ? std::string::npos : slash - i));
  return !v.empty();
}

std::string keyword(const std::string& card) { return trim(card.substr(0, 8)); }

uint64_t be_load(const unsigned char* p, int nbytes) {
  uint64_t v = 0;
  for (int i = 0; i < nbytes; ++i) v = (v << 8) | p[i];
  return v;
}

void be_store(unsigned char* p, uint64_t v, int nbytes) {
  for (int i = nbytes - 1; i >= 0; --i) {
    p[i] = (unsigned char)(v & 0xFF);
    v >>= 8;
  }
}

double decode(const unsigned char* p, int bitpix) {
  switch (bitpix) {
    case 8: return (double)p[0];
    case 16: return (double)(int16_t)be_load(p, 2);
    case 32: return (double)(int32_t)be_load(p, 4);
    case 64: return (double)(int64_t)be_load(p, 8);
    case -32: {
      uint32_t u = (uint32_t)be_load(p, 4);
      float f;
      memcpy(&f, &u, 4);
      return (double)f;
    }
    default: {
      uint64_t u = be_load(p, 8);
      double d;
      memcpy(&d, &u, 8);
      return d;
    }
  }
}

void encode(unsigned char* p, double v, int bitpix) {
  switch (bitpix) {
    case 8: p[0] = (unsigned char)std::lround(v); break;
    case 16: be_store(p, (uint16_t)(int16_t)std::lround(v), 2); break;
    case 32: be_store(p, (uint32_t)(int32_t)std::lround(v), 4); break;
    case 64: be_store(p, (uint64_t)(int64_t)std::llround(v), 8); break;
    case -32: {
      float f = (float)v;
      uint32_t u;
      memcpy(&u, &f, 4);
      be_store(p, u, 4);
      break;
    }
    default: {
      uint64_t u;
      memcpy(&u, &v, 8);
      be_store(p, u, 8);
    }
  }
}

bool valid_bitpix(int b) { return b == 8 || b == 16 || b == 32 || b == 64 || b == -32 || b == -64; }

std::string fmt_int_card(const std::string& key, int64_t v) {
  char buf[96];
  snprintf(buf, sizeof buf, "%-8s= %20lld", key.c_str(), (long long)v);
  return card80(buf);
}

std::string fmt_logical_card(const std::string& key, bool v) {
  char buf[96];
  snprintf(buf, sizeof buf, "%-8s= %20s", key.c_str(), v ? "T" : "F");
  return card80(buf);
}
}  // namespace

std::string card80(const std::string& s) {
  std::string c = s.substr(0, std::min(s.size(), kCard));
  c.resize(kCard, ' ');
  return c;
}

bool card_string(const std::vector<std::string>& cards, const std::string& key, std::string& out) {
  for (const auto& c : cards)
    if (keyword(c) == key && card_value(c, out)) return true;
  return false;
}

bool card_double(const std::vector<std::string>& cards, const std::string& key, double& out) {
  std::string v;
  if (!card_string(cards, key, v)) return false;
  for (auto& ch : v)
    if (ch == 'D' || ch == 'd') ch = 'E';  // FITS allows Fortran exponents
  char* end = nullptr;
  out = strtod(v.c_str(), &end);
  return end != v.c_str();
}

bool card_int(const std::vector<std::string>& cards, const std::string& key, int64_t& out) {
  double d;
  if (!card_double(cards, key, d)) return false;
  out = (int64_t)d;
  return (double)out == d;
}

bool FitsFile::open(const std::string& path, std::string& err) {
  path_ = path;
  hdus_.clear();
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  bytes_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  uint64_t off = 0;
  while (off + kBlock <= bytes_.size()) {
    Hdu h;
    bool end = false;
    uint64_t p = off;
    while (!end) {
      if (p + kBlock > bytes_.size()) {
        err = path + ": header without END";
        return false;
      }
      for (size_t k = 0; k < kBlock / kCard; ++k) {
        std::string card(reinterpret_cast<const char*>(&bytes_[p + k * kCard]), kCard);
        if (keyword(card) == "END") {
          end = true;
          break;
        }
        h.cards.push_back(card);
      }
      p += kBlock;
    }
    int64_t bitpix = 0, naxis = 0;
    if (!card_int(h.cards, "BITPIX", bitpix) || !valid_bitpix((int)bitpix) || !card_int(h.cards, "NAXIS", naxis) ||
        naxis < 0 || naxis > 999) {
      if (hdus_.empty()) {
        err = path + ": not a FITS image (BITPIX/NAXIS)";
        return false;
      }
      break;  // trailing garbage after a valid HDU
    }
    h.bitpix = (int)bitpix;
    uint64_t nelem = naxis ? 1 : 0;
    for (int64_t i = 1; i <= naxis; ++i) {
      int64_t ni = 0;
      if (!card_int(h.cards, "NAXIS" + std::to_string(i), ni) || ni < 0) {
        err = path + ": missing NAXIS" + std::to_string(i);
        return false;
      }
      h.naxes.push_back(ni);
      nelem *= (uint64_t)ni;
    }
    int64_t pcount = 0, gcount = 1;
    card_int(h.cards, "PCOUNT", pcount);
    card_int(h.cards, "GCOUNT", gcount);
    card_double(h.cards, "BSCALE", h.bscale);
    card_double(h.cards, "BZERO", h.bzero);
    h.nelem = nelem;
    h.data_offset = p;
    const uint64_t dbytes = (uint64_t)(std::abs(h.bitpix) / 8) * (uint64_t)gcount * (pcount + nelem);
    if (p + dbytes > bytes_.size()) {
      err = path + ": truncated data unit";
      return false;
    }
    hdus_.push_back(std::move(h));
    off = p + (dbytes + kBlock - 1) / kBlock * kBlock;
  }
  if (hdus_.empty()) {
    err = path + ": no HDU";
    return false;
  }
  return true;
}

bool FitsFile::read(size_t hdu, uint64_t first, uint64_t count, double* out, std::string& err) const {
  if (hdu >= hdus_.size()) {
    err = path_ + ": HDU " + std::to_string(hdu + 1) + " does not exist";
    return false;
  }
  const Hdu& h = hdus_[hdu];
  if (first < 1 || first - 1 + count > h.nelem) {
    err = path_ + ": read past the end of HDU " + std::to_string(hdu + 1);
    return false;
  }
  const int nb = std::abs(h.bitpix) / 8;
  const unsigned char* p = bytes_.data() + h.data_offset + (first - 1) * nb;
  for (uint64_t i = 0; i < count; ++i) out[i] = h.bzero + h.bscale * decode(p + i * nb, h.bitpix);
  return true;
}

std::string FitsFile::header_string(size_t hdu) const {
  std::string s;
  for (const auto& c : hdus_[hdu].cards) s += c;
  s += card80("END");
  return s;
}

bool write_file(const std::string& path_in, const std::vector<HduOut>& hdus, std::string& err) {
  std::string path = path_in;
  if (!path.empty() && path[0] == '!') path = path.substr(1);
  std::vector<unsigned char> out;
  for (size_t i = 0; i < hdus.size(); ++i) {
    const HduOut& h = hdus[i];
    if (!valid_bitpix(h.bitpix)) {
      err = "write " + path + ": bad BITPIX";
      return false;
    }
    std::vector<std::string> cards;
    if (i == 0)
      cards.push_back(fmt_logical_card("SIMPLE", true));
    else
      cards.push_back(card80("XTENSION= 'IMAGE   '           / IMAGE extension"));
    cards.push_back(fmt_int_card("BITPIX", h.bitpix));
    cards.push_back(fmt_int_card("NAXIS", (int64_t)h.naxes.size()));
    uint64_t nelem = h.naxes.empty() ? 0 : 1;
    for (size_t a = 0; a < h.naxes.size(); ++a) {
      cards.push_back(fmt_int_card("NAXIS" + std::to_string(a + 1), h.naxes[a]));
      nelem *= (uint64_t)h.naxes[a];
    }
    if (i == 0)
      cards.push_back(fmt_logical_card("EXTEND", true));
    else {
      cards.push_back(fmt_int_card("PCOUNT", 0));
      cards.push_back(fmt_int_card("GCOUNT", 1));
    }
    for (const auto& c : h.extra_cards) cards.push_back(card80(c));
    cards.push_back(card80("END"));
    size_t hstart = out.size();
    for (const auto& c : cards) out.insert(out.end(), c.begin(), c.end());
    out.resize(hstart + (out.size() - hstart + kBlock - 1) / kBlock * kBlock, ' ');
    const int nb = std::abs(h.bitpix) / 8;
    size_t dstart = out.size();
    out.resize(dstart + nelem * nb, 0);
    const uint64_t nd = std::min<uint64_t>(h.ndata, nelem);
    for (uint64_t k = 0; k < nd; ++k) encode(&out[dstart + k * nb], h.data[k], h.bitpix);
    out.resize(dstart + (nelem * nb + kBlock - 1) / kBlock * kBlock, 0);
  }
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    err = "cannot create " + path;
    return false;
  }
  f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size());
  if (!f) {
    err = "write failed: " + path;
    return false;
  }
  return true;
}

}  // namespace hcgio
