// parseConfig / allocDataHost / readdata with the signatures of
// fxkernel/src/common.h:41-64, implemented over the gxcorr C-ABI.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../gxcorr.h"

// fxkernel/src/common.cpp:54-114
inline void parseConfig(char *configfilename, int &nbit, int &nPol, bool &iscomplex, int &nchan, int &nant,
                        double &lo, double &bandwidth, int &numffts, std::vector<std::string> &antenna,
                        std::vector<std::string> &antFiles, double ***delays, double **antfileoffsets) {
  gxc_conffile cf;
  gxc_status s = gxc_parse_config(configfilename, &cf);
  if (s != GXC_OK) {
    std::cerr << "Error: could not parse " << configfilename << std::endl;
    std::exit(1);
  }
  nbit = cf.cfg.nbit;
  nPol = cf.cfg.npol;
  iscomplex = cf.cfg.is_complex != 0;
  nchan = cf.cfg.nchan;
  nant = cf.cfg.nant;
  lo = cf.cfg.lo;
  bandwidth = cf.cfg.bw;
  numffts = cf.cfg.numffts;
  *delays = new double *[nant > 0 ? nant : 1];
  *antfileoffsets = new double[nant > 0 ? nant : 1];
  for (int i = 0; i < nant; i++) {
    (*delays)[i] = new double[3];
    for (int k = 0; k < 3; k++) (*delays)[i][k] = cf.delays[3 * i + k];
    (*antfileoffsets)[i] = cf.offsets[i];
    if (i < cf.nant_read) {
      antenna.push_back(cf.names[i]);
      antFiles.push_back(cf.files[i]);
    }
  }
  gxc_conffile_free(&cf);
}

// fxkernel/src/common.cpp:3-36; station buffers as numffts+1 FFTs.
inline void allocDataHost(uint8_t ***data, int numantenna, int numchannels, int numffts, int nbit, int nPol,
                          bool iscomplex, int &subintbytes) {
  gxc_config c;
  gxc_config_init(&c);
  c.nant = numantenna;
  c.nchan = numchannels;
  c.numffts = numffts;
  c.nbit = nbit;
  c.npol = nPol;
  c.is_complex = iscomplex ? 1 : 0;
  c.bw = 1.0;
  const size_t bytes = gxc_station_bytes(&c);
  if (bytes == 0 || bytes > 0x7fffffffu) {
    std::cerr << "Unable to size station buffers for this configuration. Quitting" << std::endl;
    std::exit(1);
  }
  subintbytes = (int)bytes;
  std::cout << "Allocating " << subintbytes / 1024 / 1024 << " MB per antenna per subint" << std::endl;
  std::cout << "          " << (size_t)subintbytes * numantenna / 1024 / 1024 << " MB total" << std::endl;
  *data = new uint8_t *[numantenna];
  for (int a = 0; a < numantenna; a++) {
    (*data)[a] = new (std::nothrow) uint8_t[bytes];
    if (!(*data)[a]) {
      std::cerr << "Unable to allocate " << subintbytes << " bytes. Quitting" << std::endl;
      std::exit(1);
    }
  }
}

// fxkernel/src/common.cpp:38-52
inline int readdata(int bytestoread, std::vector<std::ifstream *> &antStream, uint8_t **inputdata) {
  for (size_t i = 0; i < antStream.size(); i++) {
    antStream[i]->read((char *)inputdata[i], bytestoread);
    if (!*(antStream[i])) {
      if (antStream[i]->eof()) return 2;
      std::cerr << "Error: Problem reading data (Read " << antStream[i]->gcount() << " bytes)" << std::endl;
      return 1;
    }
  }
  return 0;
}
