// .conf reader: the keyword format of fxkernel/src/common.cpp:54-114.
//
// Same keywords (COMPLEX, NBIT, NPOL, NCHAN, LO, BANDWIDTH, NUMFFTS, NANT,
// case-insensitive), same defaults (NPOL 2, COMPLEX 0, NBIT 2), NANT followed
// by NANT lines "name file d0 d1 d2 offset", unknown keywords reported on
// stderr and skipped, an unparsable (empty) line is an error.  Where the
// reference exits the process, this returns GXC_ERR_INVALID_ARG.
#include <strings.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "../../include/gxcorr.h"

namespace {
thread_local std::string g_conf_error;

char *dupstr(const std::string &s) {
  char *p = static_cast<char *>(std::malloc(s.size() + 1));
  if (p) std::memcpy(p, s.c_str(), s.size() + 1);
  return p;
}
}  // namespace

extern "C" {

void gxc_conffile_free(gxc_conffile *cf) {
  if (!cf) return;
  for (int i = 0; i < cf->cfg.nant; i++) {
    if (cf->names) std::free(cf->names[i]);
    if (cf->files) std::free(cf->files[i]);
  }
  std::free(cf->names);
  std::free(cf->files);
  std::free(cf->delays);
  std::free(cf->offsets);
  cf->names = cf->files = nullptr;
  cf->delays = cf->offsets = nullptr;
  cf->nant_read = 0;
}

gxc_status gxc_parse_config(const char *path, gxc_conffile *out) {
  if (!path || !out) return GXC_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  gxc_config_init(&out->cfg);
  std::ifstream fconfig(path);
  if (!fconfig) return GXC_ERR_IO;
  std::string line;
  int anttoread = 0, iant = 0;
  bool iscomplex = false;
  while (std::getline(fconfig, line)) {
    std::istringstream iss(line);
    std::string keyword;
    if (!(iss >> keyword)) {
      std::cerr << "Error: Could not parse \"" << line << "\"" << std::endl;
      gxc_conffile_free(out);
      return GXC_ERR_INVALID_ARG;
    }
    if (anttoread) {
      std::string thisfile;
      iss >> thisfile;
      out->names[iant] = dupstr(keyword);
      out->files[iant] = dupstr(thisfile);
      for (int i = 0; i < 3; i++) iss >> out->delays[3 * iant + i];
      iss >> out->offsets[iant];
      iant++;
      anttoread--;
      out->nant_read = iant;
    } else if (strcasecmp(keyword.c_str(), "COMPLEX") == 0) {
      iss >> iscomplex;
      out->cfg.is_complex = iscomplex ? 1 : 0;
    } else if (strcasecmp(keyword.c_str(), "NBIT") == 0) {
      iss >> out->cfg.nbit;
    } else if (strcasecmp(keyword.c_str(), "NPOL") == 0) {
      iss >> out->cfg.npol;
    } else if (strcasecmp(keyword.c_str(), "NCHAN") == 0) {
      iss >> out->cfg.nchan;
    } else if (strcasecmp(keyword.c_str(), "LO") == 0) {
      iss >> out->cfg.lo;
    } else if (strcasecmp(keyword.c_str(), "BANDWIDTH") == 0) {
      iss >> out->cfg.bw;
    } else if (strcasecmp(keyword.c_str(), "NUMFFTS") == 0) {
      iss >> out->cfg.numffts;
    } else if (strcasecmp(keyword.c_str(), "NANT") == 0) {
      int nant = 0;
      iss >> nant;
      if (nant < 0 || nant > 1 << 20) {
        gxc_conffile_free(out);
        return GXC_ERR_INVALID_ARG;
      }
      gxc_conffile_free(out);
      out->cfg.nant = nant;
      out->names = static_cast<char **>(std::calloc(nant > 0 ? nant : 1, sizeof(char *)));
      out->files = static_cast<char **>(std::calloc(nant > 0 ? nant : 1, sizeof(char *)));
      out->delays = static_cast<double *>(std::calloc(3 * (size_t)(nant > 0 ? nant : 1), sizeof(double)));
      out->offsets = static_cast<double *>(std::calloc(nant > 0 ? nant : 1, sizeof(double)));
      if (!out->names || !out->files || !out->delays || !out->offsets) {
        gxc_conffile_free(out);
        return GXC_ERR_OUT_OF_MEMORY;
      }
      anttoread = nant;
      iant = 0;
    } else {
      std::cerr << "Error: Unknown keyword \"" << keyword << "\"" << std::endl;
    }
  }
  return GXC_OK;
}

}  // extern "C"
