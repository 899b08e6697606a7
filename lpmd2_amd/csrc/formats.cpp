// Configuration files of LPMD on host arrays: the lpmd / zlp format of
// plugins/lpmd.cc (reader for LPMD 1.0, LPMD 2.0 text and compressed, and the
// legacy ZLP 1 stream; writer for LPMD 2.0 text and compressed), the xyz
// writer of plugins/xyz.cc, and SimulationEngine::Dump / Restore
// (liblpmd/simulation/simulationbuilder.cc:195-293).  No GPU is involved:
// lpmdx_snapshot_take (engine.hip) brings a system's state to the host
// asynchronously, and these routines format it.
//
// Numbers are written as the reference's iostreams write them: std::fixed is
// printf's %f (six decimals, or xyz's precout), the default float format is
// %g with six significant digits.  Cells are orthogonal: fractional
// coordinates are x / |cell row| (OrthogonalCell::Fractional,
// liblpmd/lpmd/orthogonalcell.h:90-94).
#include <zlib.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.hpp"

using namespace lpmdx;

namespace {

// ElemSym of liblpmd/support/elements.cc:9-18: the symbol each Z is written as
const char* const kElemSym[119] = {
    "e",   "H",   "He",  "Li",  "Be",  "B",   "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",  "Al",  "Si",
    "P",   "S",   "Cl",  "Ar",  "K",   "Ca",  "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co",  "Ni",  "Cu",
    "Zn",  "Ga",  "Ge",  "As",  "Se",  "Br",  "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo",  "Tc",  "Ru",
    "Rh",  "Pd",  "Ag",  "Cd",  "In",  "Sn",  "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La",  "Ce",  "Pr",
    "Nd",  "Pm",  "Sm",  "Eu",  "Gd",  "Tb",  "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",  "Ta",  "W",
    "Re",  "Os",  "Ir",  "Pt",  "Au",  "Hg",  "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr",  "Ra",  "Ac",
    "Th",  "Pa",  "U",   "Np",  "Pu",  "Am",  "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No",  "Lr",  "Rf",
    "Db",  "Sg",  "Bh",  "Hs",  "Mt",  "Uun", "Uuu", "Uub", "Uut", "Uuq", "Uup", "Uuh", "Uus", "Uuo"};

// ElemNum (liblpmd/lpmd/elements.h:16-25): the first exact match, else 0
int elem_num(const std::string& s) {
  for (int i = 0; i < 119; ++i)
    if (s == kElemSym[i])
      return i;
  return 0;
}

const char* elem_sym(int z) {
  if (z < 0 || z > 118)
    fail(LPMDX_EINVAL, "atomic number " + std::to_string(z) + " has no element symbol");
  return kElemSym[z];
}

// StringSplit(line, ' ') (liblpmd/support/util.cc:32-52 after
// RemoveUnnecessarySpaces): words separated by spaces, tabs, CR or LF
bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; }

std::vector<std::string> split(const std::string& line) {
  std::vector<std::string> w;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && is_space(line[i]))
      ++i;
    size_t j = i;
    while (j < line.size() && !is_space(line[j]))
      ++j;
    if (j > i)
      w.push_back(line.substr(i, j - i));
    i = j;
  }
  return w;
}

// atof: the longest numeric prefix, 0 when there is none
double num(const std::string& s) { return strtod(s.c_str(), nullptr); }

void put_fixed(std::string& o, double v, int prec = 6) {
  char b[512];
  snprintf(b, sizeof b, "%.*f", prec, v);
  o += b;
}

void put_g(std::string& o, double v) {
  char b[64];
  snprintf(b, sizeof b, "%g", v);
  o += b;
}

// |cell row q| (Vector::Module)
double row_len(const double* cell, int q) {
  const double* r = cell + 3 * q;
  return sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

void check_orthogonal(const double* cell) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      if (a != b && cell[3 * a + b] != 0.0)
        fail(LPMDX_EINVAL, "only orthogonal cells are written (a cell row has an off-diagonal entry)");
  for (int q = 0; q < 3; ++q)
    if (!(row_len(cell, q) > 0.0))
      fail(LPMDX_EINVAL, "cell edges must be positive");
}

// NonOrthogonalCell(a, b, c, alpha, beta, gamma)
// (liblpmd/lpmd/nonorthogonalcell.h:272-291), angles in degrees
void cell_from_lengths(const double p[6], double* cell) {
  const double d2r = M_PI / 180.0;
  const double ar = p[3] * d2r, br = p[4] * d2r, gr = p[5] * d2r;
  cell[0] = p[0];
  cell[1] = 0.0;
  cell[2] = 0.0;
  cell[3] = p[1] * cos(gr);
  cell[4] = p[1] * sin(gr);
  cell[5] = 0.0;
  cell[6] = p[2] * cos(br);
  const double t = (cos(ar) - cos(gr) * cos(br)) / sin(gr);
  cell[7] = p[2] * t;
  cell[8] = p[2] * sqrt(sin(br) * sin(br) - t * t);
}

struct File {
  FILE* fp = nullptr;
  std::string path;
  explicit File(const std::string& p, const char* mode) : path(p) {
    fp = fopen(p.c_str(), mode);
    if (!fp)
      fail(LPMDX_EINVAL, "cannot open " + p + ": " + strerror(errno));
  }
  ~File() {
    if (fp)
      fclose(fp);
  }
  void write(const std::string& s) {
    if (!s.empty() && fwrite(s.data(), 1, s.size(), fp) != s.size())
      fail(LPMDX_EINVAL, "write to " + path + " failed");
  }
  void write(const void* p, size_t k) {
    if (k && fwrite(p, 1, k, fp) != k)
      fail(LPMDX_EINVAL, "write to " + path + " failed");
  }
  // getline: false at end of file with nothing read
  bool line(std::string& s) {
    s.clear();
    int c;
    bool any = false;
    while ((c = fgetc(fp)) != EOF) {
      any = true;
      if (c == '\n')
        return true;
      s += (char)c;
    }
    return any;
  }
};

// A line source over a decompressed frame (the zlp reader reads lines from it)
struct Lines {
  const std::string* t = nullptr;
  size_t pos = 0;
  bool line(std::string& s) {
    s.clear();
    if (!t || pos >= t->size())
      return false;
    size_t e = t->find('\n', pos);
    if (e == std::string::npos)
      e = t->size();
    s = t->substr(pos, e - pos);
    pos = e + 1;
    return true;
  }
};

} // namespace

// ---------------------------------------------------------------- lpmd / zlp
struct lpmdx_cfgfile {
  std::unique_ptr<File> f;
  bool writing = false;
  bool zlp = false;
  int level = 0;
  long blocksize = 1024;
  int complev = 6;
  std::vector<std::string> extra;
  // reading
  int v0 = 0;                  // the legacy ZLP 1 stream's major version (LPMDFormat::v0)
  std::vector<std::string> hdr; // columns, "HDR" first (LPMDFormat::hdr)
  z_stream zs{};
  bool z_open = false;
  std::string text;            // the current compressed frame, decompressed
  // the last frame read
  long n = 0;
  int flevel = 0;
  double cell[9] = {0};
  std::vector<int> z;
  std::vector<double> s, v, a;

  ~lpmdx_cfgfile() {
    if (z_open) {
      if (writing)
        deflateEnd(&zs);
      else
        inflateEnd(&zs);
    }
  }

  // ---- writing (LPMDFormat::WriteHeader / WriteCell, plugins/lpmd.cc:525-642)
  void write_header() {
    std::string h = "LPMD 2.0 ";
    if (zlp) {
      zs.zalloc = Z_NULL;
      zs.zfree = Z_NULL;
      zs.opaque = Z_NULL;
      if (deflateInit(&zs, complev) != Z_OK)
        fail(LPMDX_EINVAL, "zlp: Compression failed");
      z_open = true;
      h += "Z\n";
    } else {
      h += "L\n";
    }
    std::vector<std::string> cols = {"SYM", "X", "Y", "Z"};
    if (level >= 1)
      cols.insert(cols.end(), {"VX", "VY", "VZ"});
    if (level >= 2)
      cols.insert(cols.end(), {"AX", "AY", "AZ"});
    for (auto& e : extra)
      cols.push_back(e);
    h += "HDR ";
    for (auto& c : cols)
      h += c + " ";
    h += "\n";
    f->write(h);
    fflush(f->fp);
  }

  void write_frame(long na, const double* c, const int* zz, const double* x, const double* vv, const double* aa) {
    check_orthogonal(c);
    std::string t = std::to_string(na) + "\n";
    for (int r = 0; r < 3; ++r) {
      if (r)
        t += " ";
      for (int q = 0; q < 3; ++q) {
        if (q)
          t += " ";
        put_fixed(t, c[3 * r + q]);
      }
    }
    t += "\n";
    const double L[3] = {row_len(c, 0), row_len(c, 1), row_len(c, 2)};
    for (long i = 0; i < na; ++i) {
      t += elem_sym(zz[i]);
      t += " ";
      for (int q = 0; q < 3; ++q) {
        if (q)
          t += " ";
        put_fixed(t, x[3 * i + q] / L[q]);
      }
      if (level >= 1)
        for (int q = 0; q < 3; ++q) {
          t += " ";
          put_fixed(t, vv ? vv[3 * i + q] : 0.0);
        }
      if (level >= 2)
        for (int q = 0; q < 3; ++q) {
          t += " ";
          put_fixed(t, aa ? aa[3 * i + q] : 0.0);
        }
      for (auto& e : extra)
        if (e == "TYPE" || e == "type")
          t += "          ATOMTYPE";
      t += "\n";
    }
    if (!zlp) {
      f->write(t);
      return;
    }
    // the frame deflated in blocksize pieces, each ending with a sync flush,
    // on the stream the header opened; then its size as one byte
    // sizeof(unsigned long) and an unsigned long holding htonl(size)
    std::string cb;
    std::vector<unsigned char> out(blocksize);
    for (size_t p = 0; p < t.size(); p += blocksize) {
      const size_t k = std::min<size_t>(blocksize, t.size() - p);
      zs.avail_in = (uInt)k;
      zs.next_in = (Bytef*)(t.data() + p);
      do {
        zs.avail_out = (uInt)blocksize;
        zs.next_out = out.data();
        if (deflate(&zs, Z_SYNC_FLUSH) == Z_STREAM_ERROR)
          fail(LPMDX_EINVAL, "zlp: Compression failed");
        cb.append((const char*)out.data(), blocksize - zs.avail_out);
      } while (zs.avail_out == 0);
    }
    const unsigned char width = (unsigned char)sizeof(unsigned long);
    unsigned char len[sizeof(unsigned long)] = {0};
    const uint32_t sz = (uint32_t)cb.size();
    len[0] = (unsigned char)(sz >> 24);
    len[1] = (unsigned char)(sz >> 16);
    len[2] = (unsigned char)(sz >> 8);
    len[3] = (unsigned char)sz;
    f->write(&width, 1);
    f->write(len, width);
    f->write(cb);
    fflush(f->fp);
  }

  // ---- reading (LPMDFormat::ReadHeader / ReadCell, plugins/lpmd.cc:125-430)
  void init_inflate() {
    zs.zalloc = Z_NULL;
    zs.zfree = Z_NULL;
    zs.opaque = Z_NULL;
    if (inflateInit(&zs) != Z_OK)
      fail(LPMDX_EINVAL, "zlp: Decompression failed");
    z_open = true;
  }

  void read_header() {
    int c0 = fgetc(f->fp);
    if (c0 == 'Z') {
      // ReadHeaderZLPOne (lpmd.cc:212-226): "Z4L2P", version bytes, two flags
      unsigned char h[10];
      h[0] = 'Z';
      if (fread(&h[1], 1, 9, f->fp) != 9 || h[2] != 'L' || h[4] != 'P' || h[1] != '4' || h[3] != '2')
        fail(LPMDX_EINVAL, "zlp: Wrong header");
      v0 = h[5];
      zlp = true;
      init_inflate();
      return;
    }
    if (c0 != 'L')
      fail(LPMDX_EINVAL, "lpmd: Wrong header (not an LPMD/ZLP file)");
    std::string line;
    f->line(line);
    std::vector<std::string> w = split("L" + line);
    if (w.empty() || w[0] != "LPMD")
      fail(LPMDX_EINVAL, "lpmd: Wrong header (not an LPMD/ZLP file)");
    const std::string version = w.size() > 1 ? w[1] : "";
    zlp = false;
    if (w.size() > 2) {
      if (w[2] == "Z")
        zlp = true;
      else if (w[2] != "L")
        fail(LPMDX_EINVAL, "lpmd: Wrong compression flag (corrupted LPMD/ZLP file?)");
    }
    if (version == "1.0") {
      // ReadHeaderLPMDOne (lpmd.cc:171-210): the third line's word count
      // gives the columns; the file position is kept
      const long where = ftell(f->fp);
      std::string info;
      f->line(info);
      f->line(info);
      f->line(info);
      const size_t k = split(info).size();
      fseek(f->fp, where, SEEK_SET);
      hdr = {"HDR", "SYM", "X", "Y", "Z"};
      if (k == 4) {
        level = 0;
      } else if (k == 7) {
        hdr.insert(hdr.end(), {"VX", "VY", "VZ"});
        level = 1;
      } else if (k == 10) {
        // force columns are named FX FY FZ there, which no reader column
        // matches: the accelerations stay zero
        hdr.insert(hdr.end(), {"VX", "VY", "VZ", "FX", "FY", "FZ"});
        level = 2;
      } else {
        fail(LPMDX_EINVAL, "lpmd: File " + f->path + " not have a apropiate 1.0 version");
      }
    } else if (version == "2.0") {
      std::string h;
      f->line(h);
      std::vector<std::string> words = split(h);
      if (words.empty() || words[0] != "HDR")
        fail(LPMDX_EINVAL, "lpmd: File " + f->path + " doesn't seem to be in LPMD 2.0 format (wrong HDR)");
      hdr = words;
    } else {
      fail(LPMDX_EINVAL, "lpmd: Format version " + version + " in file " + f->path + " is not supported.");
    }
    if (zlp)
      init_inflate();
  }

  // One compressed frame: a byte giving the width of the size field, the
  // size (the first four bytes, network order), the deflated bytes; false at
  // the end of the file.
  bool inflate_frame() {
    int w = fgetc(f->fp);
    if (w == EOF)
      return false;
    unsigned char len[16] = {0};
    if (w < 4 || w > 16 || fread(len, 1, w, f->fp) != (size_t)w)
      fail(LPMDX_EINVAL, "zlp: truncated frame size");
    const uint32_t sz = ((uint32_t)len[0] << 24) | ((uint32_t)len[1] << 16) | ((uint32_t)len[2] << 8) | len[3];
    std::vector<unsigned char> in(sz), out(blocksize);
    if (sz && fread(in.data(), 1, sz, f->fp) != sz)
      fail(LPMDX_EINVAL, "zlp: truncated frame");
    text.clear();
    zs.avail_in = sz;
    zs.next_in = in.data();
    for (;;) {
      const uInt in0 = zs.avail_in;
      zs.avail_out = (uInt)blocksize;
      zs.next_out = out.data();
      const int rc = inflate(&zs, Z_SYNC_FLUSH);
      if (rc != Z_OK && rc != Z_STREAM_END && rc != Z_BUF_ERROR)
        fail(LPMDX_EINVAL, "zlp: Decompression failed");
      const size_t got = blocksize - zs.avail_out;
      text.append((const char*)out.data(), got);
      if (rc == Z_STREAM_END || zs.avail_out != 0)
        break; // the frame's input is used up (or the stream ended)
      if (got == 0 && zs.avail_in == in0)
        fail(LPMDX_EINVAL, "zlp: Decompression failed");
    }
    return true;
  }

  void resize(long na) {
    n = na;
    z.assign(na, 0);
    s.assign(3 * na, 0.0);
    v.assign(3 * na, 0.0);
    a.assign(3 * na, 0.0);
  }

  // the next frame into (n, cell, z, s, v, a); false at the end
  bool read_frame() {
    if (zlp && !inflate_frame())
      return false;
    if (v0 == 1) {
      // the legacy ZLP 1 frame (lpmd.cc:267-309): level, atoms, nine cell
      // numbers, then per atom the symbol, scaled position, velocity and
      // acceleration as the level has them, whitespace separated
      std::vector<std::string> w = split(text);
      size_t p = 0;
      auto next = [&]() -> const std::string& {
        if (p >= w.size())
          fail(LPMDX_EINVAL, "zlp: truncated frame");
        return w[p++];
      };
      const int lvl = atoi(next().c_str());
      const long na = atol(next().c_str());
      if (na < 0)
        fail(LPMDX_EINVAL, "zlp: negative atom count");
      for (int k = 0; k < 9; ++k)
        cell[k] = num(next());
      resize(na);
      flevel = lvl;
      for (long i = 0; i < na; ++i) {
        z[i] = elem_num(next());
        for (int q = 0; q < 3; ++q)
          s[3 * i + q] = num(next());
        if (lvl > 0)
          for (int q = 0; q < 3; ++q)
            v[3 * i + q] = num(next());
        if (lvl > 1)
          for (int q = 0; q < 3; ++q)
            a[3 * i + q] = num(next());
      }
      return true;
    }
    Lines src;
    src.t = &text;
    auto line = [&](std::string& l) { return zlp ? src.line(l) : f->line(l); };
    std::string l;
    if (!line(l))
      return false;
    std::vector<std::string> w = split(l);
    if (w.empty())
      return false;
    const long na = atol(w[0].c_str());
    line(l);
    w = split(l);
    if (w.size() == 9) {
      for (int k = 0; k < 9; ++k)
        cell[k] = num(w[k]);
    } else if (w.size() == 6) {
      double p6[6];
      for (int k = 0; k < 6; ++k)
        p6[k] = num(w[k]);
      cell_from_lengths(p6, cell);
    } else {
      fail(LPMDX_EINVAL, "lpmd: Error ocurred when reading the base vectors, file \"" + f->path + "\"");
    }
    resize(std::max<long>(na, 0));
    flevel = level;
    for (long i = 0; i < n; ++i) {
      if (!line(l))
        l.clear();
      w = split(l);
      if (w.empty())
        fail(LPMDX_EINVAL, "lpmd: Error ocurred, the atom file not have elements!");
      if (w.size() + 1 < hdr.size())
        fail(LPMDX_EINVAL, "lpmd: an atom line has fewer columns than the header in " + f->path);
      // columns by the header's names (lpmd.cc:387-414); RGB and C are colours,
      // which this engine does not keep
      for (size_t k = 1; k < hdr.size(); ++k) {
        const std::string& h = hdr[k];
        const std::string& x = w[k - 1];
        if (h == "SYM")
          z[i] = elem_num(x);
        else if (h == "X")
          s[3 * i] = num(x);
        else if (h == "Y")
          s[3 * i + 1] = num(x);
        else if (h == "Z")
          s[3 * i + 2] = num(x);
        else if (h == "VX")
          v[3 * i] = num(x);
        else if (h == "VY")
          v[3 * i + 1] = num(x);
        else if (h == "VZ")
          v[3 * i + 2] = num(x);
        else if (h == "AX")
          a[3 * i] = num(x);
        else if (h == "AY")
          a[3 * i + 1] = num(x);
        else if (h == "AZ")
          a[3 * i + 2] = num(x);
      }
    }
    return true;
  }
};

namespace {

void need_ptr(const void* p, const char* what) {
  if (!p)
    fail(LPMDX_EINVAL, std::string(what) + " is NULL");
}

// the lpmd plugin's file type: `type`, overridden by the file name's last
// four characters "lpmd" or ".zlp" (LPMDFormat::LPMDFormat, lpmd.cc:37-44)
bool zlp_type(const std::string& path, const std::string& type) {
  bool z = type == "zlp";
  if (path.size() >= 4) {
    const std::string q = path.substr(path.size() - 4);
    if (q == "lpmd")
      z = false;
    else if (q == ".zlp")
      z = true;
  }
  return z;
}

} // namespace

extern "C" {

int lpmdx_lpmd_open_write(lpmdx_cfgfile** out, const char* path, const char* args) {
  return guard([&] {
    need_ptr(out, "out");
    need_ptr(path, "path");
    *out = nullptr;
    Keywords k(args, {{"level", "0"}, {"extra", ""}, {"blocksize", "1024"}, {"compression", "6"}, {"type", "lpmd"}});
    auto F = std::make_unique<lpmdx_cfgfile>();
    F->writing = true;
    F->level = k.i("level");
    F->blocksize = k.i("blocksize");
    F->complev = k.i("compression");
    if (F->blocksize <= 0)
      fail(LPMDX_EINVAL, "lpmd: blocksize must be positive");
    const std::string ex = k.s("extra");
    size_t p = 0;
    while (p <= ex.size() && !ex.empty()) {
      size_t e = ex.find(',', p);
      if (e == std::string::npos)
        e = ex.size();
      const std::string it = ex.substr(p, e - p);
      if (it == "TYPE" || it == "type")
        F->extra.push_back(it);
      else
        fail(LPMDX_EINVAL, "lpmd: extra column " + it + " is not written (atom colours are not kept)");
      p = e + 1;
    }
    F->zlp = zlp_type(path, k.s("type"));
    F->f = std::make_unique<File>(path, "wb");
    F->write_header();
    *out = F.release();
  });
}

int lpmdx_lpmd_write(lpmdx_cfgfile* F, long n, const double cell[9], const int* z, const double* x, const double* v,
                     const double* a) {
  return guard([&] {
    need_ptr(F, "file");
    if (!F->writing)
      fail(LPMDX_EINVAL, "lpmd: file was opened for reading");
    if (n < 0)
      fail(LPMDX_EINVAL, "negative atom count");
    need_ptr(cell, "cell");
    if (n > 0) {
      need_ptr(z, "z");
      need_ptr(x, "x");
    }
    F->write_frame(n, cell, z, x, v, a);
  });
}

int lpmdx_lpmd_open_read(lpmdx_cfgfile** out, const char* path, int* level) {
  return guard([&] {
    need_ptr(out, "out");
    need_ptr(path, "path");
    *out = nullptr;
    auto F = std::make_unique<lpmdx_cfgfile>();
    F->f = std::make_unique<File>(path, "rb");
    F->read_header();
    if (level)
      *level = F->v0 == 1 ? -1 : F->level;
    *out = F.release();
  });
}

int lpmdx_lpmd_next(lpmdx_cfgfile* F, long* n) {
  return guard([&] {
    need_ptr(F, "file");
    need_ptr(n, "n");
    if (F->writing)
      fail(LPMDX_EINVAL, "lpmd: file was opened for writing");
    *n = F->read_frame() ? F->n : -1;
  });
}

int lpmdx_lpmd_frame(lpmdx_cfgfile* F, double cell[9], int* z, double* s, double* v, double* a, int* level) {
  return guard([&] {
    need_ptr(F, "file");
    if (cell)
      memcpy(cell, F->cell, sizeof F->cell);
    const size_t m = (size_t)F->n;
    if (z && m)
      memcpy(z, F->z.data(), m * sizeof(int));
    if (s && m)
      memcpy(s, F->s.data(), 3 * m * sizeof(double));
    if (v && m)
      memcpy(v, F->v.data(), 3 * m * sizeof(double));
    if (a && m)
      memcpy(a, F->a.data(), 3 * m * sizeof(double));
    if (level)
      *level = F->flevel;
  });
}

int lpmdx_lpmd_close(lpmdx_cfgfile* F) {
  return guard([&] { delete F; });
}

// ---------------------------------------------------------------- xyz
int lpmdx_xyz_write(const char* path, int append, long n, const double cell[9], const int* z, const double* x,
                    const double* v, const double* a, const char* args) {
  return guard([&] {
    need_ptr(path, "path");
    need_ptr(cell, "cell");
    if (n < 0)
      fail(LPMDX_EINVAL, "negative atom count");
    if (n > 0) {
      need_ptr(z, "z");
      need_ptr(x, "x");
    }
    // XYZFormat's keywords (plugins/xyz.cc:17-26)
    Keywords k(args, {{"level", "0"}, {"each", "1"}, {"coords", "positive"}, {"inside", "false"},
                      {"external", "consider"}, {"zerocm", "false"}, {"replacecell", "false"}, {"precout", "8"}});
    const int level = k.i("level");
    const int prec = k.i("precout");
    const bool inside = k.s("inside") == "true";
    if (inside)
      check_orthogonal(cell);
    if ((level >= 1 && !v) || (level >= 2 && !a))
      fail(LPMDX_EINVAL, "xyz: the level asks for velocities or accelerations that were not given");
    // XYZFormat::WriteCell (plugins/xyz.cc:226-249); FittedInside with inside
    // true (orthogonalcell.h:102-114), on the written copy only
    std::string t = std::to_string(n) + "\n\n";
    const double L[3] = {row_len(cell, 0), row_len(cell, 1), row_len(cell, 2)};
    auto vec = [&](const double* p, bool fit) {
      for (int q = 0; q < 3; ++q) {
        double u = p[q];
        if (fit && (u < 0.0 || u > L[q]))
          u -= L[q] * floor(u / L[q]);
        if (q)
          t += " ";
        put_fixed(t, u, prec);
      }
    };
    for (long i = 0; i < n; ++i) {
      t += elem_sym(z[i]);
      t += " ";
      vec(x + 3 * i, inside);
      if (level >= 1) {
        t += " ";
        vec(v + 3 * i, false);
      }
      if (level >= 2) {
        t += " ";
        vec(a + 3 * i, false);
      }
      t += "\n";
    }
    File f(path, append ? "ab" : "wb");
    f.write(t);
  });
}

// ---------------------------------------------------------------- Dump / Restore
int lpmdx_dump_write(const char* path, long step, const double cell[9], long n, const int* z, const double* x,
                     const double* v, const double* a, const int* flags) {
  return guard([&] {
    need_ptr(path, "path");
    need_ptr(cell, "cell");
    if (n < 0)
      fail(LPMDX_EINVAL, "negative atom count");
    if (n > 0) {
      need_ptr(z, "z");
      need_ptr(x, "x");
      need_ptr(v, "v");
      need_ptr(a, "a");
    }
    // SimulationEngine::Dump (simulationbuilder.cc:195-239): the step, the
    // cell rows, the atoms (Z, position, velocity, acceleration), one colour
    // line and one tag line per atom, in the stream's default %g format
    std::string t = std::to_string(step) + "\n";
    for (int r = 0; r < 3; ++r) {
      for (int q = 0; q < 3; ++q) {
        put_g(t, cell[3 * r + q]);
        t += " ";
      }
      t += "\n";
    }
    t += std::to_string(n) + "\n";
    for (long i = 0; i < n; ++i) {
      t += std::to_string(z[i]) + " ";
      for (const double* p : {x, v, a})
        for (int q = 0; q < 3; ++q) {
          put_g(t, p[3 * i + q]);
          t += " ";
        }
      t += "\n";
    }
    for (long i = 0; i < n; ++i)
      t += "false \n";
    // the tags this engine keeps: fixedvel (any value reads as set), fixedpos
    // "true", or fixedpos with another value when only tagged
    for (long i = 0; i < n; ++i) {
      const int f = flags ? flags[i] : 0;
      std::vector<std::pair<const char*, const char*>> tg;
      if (f & LPMDX_FIX_VEL)
        tg.push_back({"fixedvel", "true"});
      if (f & LPMDX_FIX_POS)
        tg.push_back({"fixedpos", "true"});
      else if ((f & LPMDX_FIX_TAGGED) && !(f & LPMDX_FIX_VEL))
        tg.push_back({"fixedpos", "false"});
      t += std::to_string(tg.size()) + " ";
      for (auto& p : tg)
        t += std::string(p.first) + " " + p.second + " ";
      t += "\n";
    }
    File f(path, "wb");
    f.write(t);
  });
}

int lpmdx_dump_read(const char* path, long* step, double cell[9], long* n, long cap, int* z, double* x, double* v,
                    double* a, int* flags) {
  return guard([&] {
    need_ptr(path, "path");
    need_ptr(n, "n");
    File f(path, "rb");
    // SimulationEngine::Restore (simulationbuilder.cc:241-290): whitespace
    // separated fields read with >>
    std::string all;
    {
      char buf[65536];
      size_t k;
      while ((k = fread(buf, 1, sizeof buf, f.fp)) > 0)
        all.append(buf, k);
    }
    std::vector<std::string> w = split(all);
    size_t p = 0;
    auto next = [&]() -> const std::string& {
      if (p >= w.size())
        fail(LPMDX_EINVAL, "restore: " + std::string(path) + " ends early");
      return w[p++];
    };
    const long st = atol(next().c_str());
    double c[9];
    for (int k = 0; k < 9; ++k)
      c[k] = num(next());
    const long na = atol(next().c_str());
    if (na < 0)
      fail(LPMDX_EINVAL, "restore: negative atom count");
    if (step)
      *step = st;
    if (cell)
      memcpy(cell, c, sizeof c);
    *n = na;
    if (cap < na)
      return; // sizes only
    std::vector<int> zz(na), ff(na, 0);
    std::vector<double> xx(3 * na), vv(3 * na), aa(3 * na);
    for (long i = 0; i < na; ++i) {
      zz[i] = atoi(next().c_str());
      for (auto* dst : {&xx, &vv, &aa})
        for (int q = 0; q < 3; ++q)
          (*dst)[3 * i + q] = num(next());
    }
    for (long i = 0; i < na; ++i)
      if (next() == "true")
        for (int q = 0; q < 3; ++q)
          next(); // a colour, which this engine does not keep
    for (long i = 0; i < na; ++i) {
      const long k = atol(next().c_str());
      for (long j = 0; j < k; ++j) {
        const std::string key = next();
        const std::string val = next();
        if (key == "fixedvel")
          ff[i] |= LPMDX_FIX_VEL | LPMDX_FIX_TAGGED;
        else if (key == "fixedpos")
          ff[i] |= LPMDX_FIX_TAGGED | (val == "true" ? LPMDX_FIX_POS : 0);
      }
    }
    if (z)
      memcpy(z, zz.data(), na * sizeof(int));
    if (x)
      memcpy(x, xx.data(), 3 * na * sizeof(double));
    if (v)
      memcpy(v, vv.data(), 3 * na * sizeof(double));
    if (a)
      memcpy(a, aa.data(), 3 * na * sizeof(double));
    if (flags)
      memcpy(flags, ff.data(), na * sizeof(int));
  });
}

} // extern "C"
