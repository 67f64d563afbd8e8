// io_format.h -- the reference's on-disk lattice data format, host side.
//
// A data file holds one record per lattice site, sites in x-slowest,
// z-fastest order over the whole system (one file, I/O grid 1 1 1):
//   binary: `count` doubles in host byte order (field.c:896-927,
//           lb_data.c:1533-1580);
//   ascii:  each value as " %22.15e" (23 characters); a newline after
//           every `per_line` values (field.c:937-996: all of a site's
//           values on one line; lb_data.c:1586-1640: one line per
//           velocity).
// Names follow io_subfile.c:186-204 ("stub-ttttttttt.001-001") and
// io_metadata.c:246-281 ("stub-metadata.001-001", a JSON document in the
// layout cJSON_Print gives io_metadata_to_json(), io_metadata.c:159-199).

#ifndef LDW_IO_FORMAT_H
#define LDW_IO_FORMAT_H

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ludwig_hip.h"

namespace ldw {
namespace io {

// Width of one ascii value, " %22.15e" (lb_data.h:45)
constexpr int ASCII_WIDTH = 23;

struct Layout {
  int ntotal[3];     // system size
  int count;         // values per site
  int per_line;      // ascii: values between newlines
};

std::string data_name(const char * stub, int timestep);
std::string metadata_name(const char * stub);

// Bytes of one site's record in the given format
int64_t record_bytes(const Layout & l, io_record_format_enum_t fmt);

// rec[site*count + c], site = (x*ny + y)*nz + z.  0 on success, otherwise
// the reason in `err`.
int write_data(const std::string & path, const Layout & l, io_record_format_enum_t fmt,
	       const double * rec, std::string & err);
int read_data(const std::string & path, const Layout & l, io_record_format_enum_t fmt,
	      double * rec, std::string & err);

// The metadata document for one quantity (options of the output side)
std::string metadata_json(const Layout & l, const io_options_t & opts, const double lmin[3],
			  const lees_edw_options_t & leo);
int write_text(const std::string & path, const std::string & text, std::string & err);

// The I/O decomposition and format are ones this library writes
int options_supported(const io_options_t & opts, std::string & err);

}  // namespace io
}  // namespace ldw

#endif
