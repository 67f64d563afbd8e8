// io_format.cpp -- the reference's on-disk lattice data format, host side
// (the layout is described in io_format.h).

#include "io_format.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>

namespace ldw {
namespace io {

std::string data_name(const char * stub, int timestep) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s-%9.9d.%3.3d-%3.3d", stub, timestep, 1, 1);
  return buf;
}

std::string metadata_name(const char * stub) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s-metadata.%3.3d-%3.3d", stub, 1, 1);
  return buf;
}

static int64_t nsites(const Layout & l) {
  return (int64_t) l.ntotal[0]*l.ntotal[1]*l.ntotal[2];
}

int64_t record_bytes(const Layout & l, io_record_format_enum_t fmt) {
  if (fmt == IO_RECORD_BINARY) return (int64_t) l.count*sizeof(double);
  return (int64_t) l.count*ASCII_WIDTH + l.count/l.per_line;
}

int options_supported(const io_options_t & o, std::string & err) {
  if (o.mode != IO_MODE_MPIIO) {
    err = "i/o mode is not mpiio";
    return -1;
  }
  if (o.iorformat != IO_RECORD_ASCII && o.iorformat != IO_RECORD_BINARY) {
    err = "record format is neither ascii nor binary";
    return -1;
  }
  if (o.iogrid[0] != 1 || o.iogrid[1] != 1 || o.iogrid[2] != 1) {
    char buf[128];
    snprintf(buf, sizeof(buf), "i/o grid %d %d %d: one GPU writes one file (1 1 1)",
	     o.iogrid[0], o.iogrid[1], o.iogrid[2]);
    err = buf;
    return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Data files                                                               */
/* ------------------------------------------------------------------------ */

static int format_ascii(double v, char * out) {
  char tmp[64];
  int n = snprintf(tmp, sizeof(tmp), " %22.15e", v);
  if (n != ASCII_WIDTH) return -1;
  memcpy(out, tmp, ASCII_WIDTH);
  return 0;
}

int write_data(const std::string & path, const Layout & l, io_record_format_enum_t fmt,
	       const double * rec, std::string & err) {
  FILE * fp = fopen(path.c_str(), "wb");
  if (fp == nullptr) {
    err = "cannot open " + path + " for writing";
    return -1;
  }
  const int64_t n = nsites(l);
  int ifail = 0;
  if (fmt == IO_RECORD_BINARY) {
    const size_t nd = (size_t) n*l.count;
    if (fwrite(rec, sizeof(double), nd, fp) != nd) ifail = -1;
  }
  else {
    // a block of sites at a time
    const int64_t rb = record_bytes(l, fmt);
    const int64_t block = 4096;
    std::vector<char> buf((size_t) (rb*block));
    for (int64_t s0 = 0; s0 < n && ifail == 0; s0 += block) {
      const int64_t s1 = (s0 + block < n) ? s0 + block : n;
      char * p = buf.data();
      for (int64_t s = s0; s < s1; s++) {
	for (int c = 0; c < l.count; c++) {
	  if (format_ascii(rec[s*l.count + c], p)) {
	    err = "value does not fit the ascii record width in " + path;
	    fclose(fp);
	    return -1;
	  }
	  p += ASCII_WIDTH;
	  if ((c + 1) % l.per_line == 0) *p++ = '\n';
	}
      }
      const size_t nb = (size_t) (p - buf.data());
      if (fwrite(buf.data(), 1, nb, fp) != nb) ifail = -1;
    }
  }
  if (fclose(fp) != 0) ifail = -1;
  if (ifail) err = "write to " + path + " failed";
  return ifail;
}

int read_data(const std::string & path, const Layout & l, io_record_format_enum_t fmt,
	      double * rec, std::string & err) {
  FILE * fp = fopen(path.c_str(), "rb");
  if (fp == nullptr) {
    err = "cannot open " + path + " for reading";
    return -1;
  }
  const int64_t n = nsites(l);
  const int64_t want = n*record_bytes(l, fmt);
  if (fseek(fp, 0, SEEK_END) != 0 || ftell(fp) != want) {
    char buf[160];
    snprintf(buf, sizeof(buf), " holds %ld bytes, not the %lld of %lld records of %lld bytes",
	     (long) ftell(fp), (long long) want, (long long) n, (long long) record_bytes(l, fmt));
    err = path + buf;
    fclose(fp);
    return -1;
  }
  rewind(fp);
  int ifail = 0;
  if (fmt == IO_RECORD_BINARY) {
    const size_t nd = (size_t) n*l.count;
    if (fread(rec, sizeof(double), nd, fp) != nd) ifail = -1;
    if (ifail) err = "read from " + path + " failed";
  }
  else {
    const int64_t rb = record_bytes(l, fmt);
    std::vector<char> buf((size_t) rb);
    for (int64_t s = 0; s < n && ifail == 0; s++) {
      if (fread(buf.data(), 1, (size_t) rb, fp) != (size_t) rb) {
	ifail = -1;
	err = "read from " + path + " failed";
	break;
      }
      const char * p = buf.data();
      for (int c = 0; c < l.count; c++) {
	char tmp[ASCII_WIDTH + 1];
	memcpy(tmp, p, ASCII_WIDTH);
	tmp[ASCII_WIDTH] = '\0';
	if (sscanf(tmp, "%le", rec + s*l.count + c) != 1) {
	  ifail = -1;
	  err = "bad ascii value in " + path;
	  break;
	}
	p += ASCII_WIDTH;
	if ((c + 1) % l.per_line == 0) p++;
      }
    }
  }
  fclose(fp);
  return ifail;
}

/* ------------------------------------------------------------------------ */
/* Metadata (formatted as cJSON_Print prints: tab indents, "key":\tvalue)   */
/* ------------------------------------------------------------------------ */

namespace {

struct Obj {
  std::vector<std::pair<std::string, std::string>> items;   // rendered values
  void add(const char * k, const std::string & v) { items.emplace_back(k, v); }
};

// cJSON's print_number (util_cJSON.c:549-583): an integral value as %d,
// else %1.15g unless that does not read back to within one ulp-scale
// (compare_double), then %1.17g
std::string num(double v) {
  char buf[64];
  const int vi = (v >= 2147483647.0) ? 2147483647 : ((v <= -2147483648.0) ? (-2147483647 - 1) : (int) v);
  if (std::isnan(v) || std::isinf(v)) snprintf(buf, sizeof(buf), "null");
  else if ((v - (double) vi) == 0.0) snprintf(buf, sizeof(buf), "%d", vi);
  else {
    snprintf(buf, sizeof(buf), "%1.15g", v);
    double test = 0.0;
    const double mx = (fabs(test = strtod(buf, nullptr)) > fabs(v)) ? fabs(test) : fabs(v);
    if (!(fabs(test - v) <= mx*DBL_EPSILON)) snprintf(buf, sizeof(buf), "%1.17g", v);
  }
  return buf;
}

std::string str(const char * s) { return std::string("\"") + s + "\""; }
std::string boolean(int b) { return b ? "true" : "false"; }

template <typename T>
std::string array3(const T v[3]) {
  return "[" + num(v[0]) + ", " + num(v[1]) + ", " + num(v[2]) + "]";
}

// An object whose values are already rendered at nesting depth `depth`
std::string render(const Obj & o, int depth) {
  std::string s = "{\n";
  for (size_t i = 0; i < o.items.size(); i++) {
    s += std::string(depth, '\t') + str(o.items[i].first.c_str()) + ":\t" + o.items[i].second;
    s += (i + 1 < o.items.size()) ? ",\n" : "\n";
  }
  s += std::string(depth - 1, '\t') + "}";
  return s;
}

}  // namespace

std::string metadata_json(const Layout & l, const io_options_t & opts, const double lmin[3],
			  const lees_edw_options_t & leo) {
  const int one[3] = {1, 1, 1};
  const int zero[3] = {0, 0, 0};
  const bool ascii = (opts.iorformat == IO_RECORD_ASCII);

  Obj options;                        // coords.c cs_options_to_json
  options.add("System size (total)", array3(l.ntotal));
  options.add("Periodic boundaries", array3(one));
  options.add("Left-end limit Lmin", array3(lmin));
  Obj le;                             // lees_edw_opts_to_json (lees_edwards_options.c:65-100)
  le.add("Number of planes", num(leo.nplanes));
  if (leo.nplanes > 0) {
    const char * type = (leo.type == LE_SHEAR_TYPE_STEADY) ? "STEADY"
      : ((leo.type == LE_SHEAR_TYPE_OSCILLATORY) ? "OSCILLATORY" : "INVALID");
    le.add("Shear type", str(type));
    if (leo.type == LE_SHEAR_TYPE_OSCILLATORY) le.add("Period (timesteps)", num(leo.period));
    le.add("Reference time", num(leo.nt0));
    le.add("Plane speed", num(leo.uy));
  }
  Obj coords;
  coords.add("options", render(options, 3));
  coords.add("lees_edwards", render(le, 3));

  Obj io;                             // io_options.c:330-350
  io.add("Mode", str("mpiio"));
  io.add("Record format", str(ascii ? "ascii" : "binary"));
  io.add("Metadata version", num((int) opts.metadata_version));
  io.add("Report", boolean(opts.report));
  io.add("Asynchronous", boolean(opts.asynchronous));
  io.add("Compression level", num(opts.compression_levl));
  io.add("I/O grid", array3(opts.iogrid));

  Obj el;                             // io_element.c:110-125
  el.add("MPI_Datatype", str(ascii ? "MPI_CHAR" : "MPI_DOUBLE"));
  el.add("Size (bytes)", num(ascii ? 1 : 8));
  el.add("Count", num((double) record_bytes(l, opts.iorformat)/(ascii ? 1 : 8)));
  el.add("Endianness", str("LITTLE_ENDIAN"));

  Obj sub;                            // io_subfile.c to_json
  sub.add("Number of files", num(1));
  sub.add("File index", num(0));
  sub.add("Topology", array3(one));
  sub.add("Coordinate", array3(zero));
  sub.add("Data ndims", num(3));
  sub.add("File size (sites)", array3(l.ntotal));
  sub.add("File offset (sites)", array3(zero));

  Obj root;
  root.add("coords", render(coords, 2));
  root.add("io_options", render(io, 2));
  root.add("io_element", render(el, 2));
  root.add("io_subfile", render(sub, 2));
  return render(root, 1);
}

int write_text(const std::string & path, const std::string & text, std::string & err) {
  FILE * fp = fopen(path.c_str(), "w");
  if (fp == nullptr) {
    err = "cannot open " + path + " for writing";
    return -1;
  }
  int ifail = (fputs(text.c_str(), fp) < 0) ? -1 : 0;
  if (fclose(fp) != 0) ifail = -1;
  if (ifail) err = "write to " + path + " failed";
  return ifail;
}

}  // namespace io
}  // namespace ldw
