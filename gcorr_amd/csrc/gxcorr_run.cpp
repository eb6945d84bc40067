// gxcorr_run <config> [vis.bin] -- the testfxkernel harness
// (fxkernel/src/testfxkernel.cpp:21-124) on the MI355X path: parse the .conf,
// read one subintegration of packed data per station, run FxKernel::process
// once, write vis.out and runtime.fxkernel.log.  Code written against
// fxkernel compiles against the facade headers unchanged.  The optional
// second argument also writes the visibilities with their row counts in the
// binary layout of gxc_save_visibilities_binary.
#include <chrono>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gxcorr/common.hpp"
#include "gxcorr/fxkernel.hpp"

using std::cerr;
using std::cout;
using std::endl;
using std::string;
using std::vector;

int main(int argc, char *argv[]) {
  if (argc < 2 || argc > 3) {
    cout << "Usage:  gxcorr_run <config> [vis.bin]\n" << endl;
    return 1;
  }
  char *configfile = argv[1];
  int subintbytes, numchannels, numantennas, numffts, nbit, nPol;
  double lo, bandwidth;
  bool iscomplex;
  u8 **inputdata;
  double **delays;
  double *antfileoffsets;
  vector<string> antennas, antFiles;
  vector<std::ifstream *> antStream;

  parseConfig(configfile, nbit, nPol, iscomplex, numchannels, numantennas, lo, bandwidth, numffts, antennas,
              antFiles, &delays, &antfileoffsets);
  cout << "Got COMPLEX " << iscomplex << endl;
  cout << "Got NBIT " << nbit << endl;
  cout << "Got NPOL " << nPol << endl;
  cout << "Got NCHAN " << numchannels << endl;
  cout << "Got LO " << lo << endl;
  cout << "Got BANDWIDTH " << bandwidth << endl;
  cout << "Got NUMFFTS " << numffts << endl;
  cout << "Got NANT " << numantennas << endl;
  for (int i = 0; i < numantennas && i < (int)antennas.size(); i++)
    cout << "  " << antennas[i] << ":" << antFiles[i] << endl;

  allocDataHost(&inputdata, numantennas, numchannels, numffts, nbit, nPol, iscomplex, subintbytes);
  for (int i = 0; i < numantennas; i++) {
    antStream.push_back(new std::ifstream(antFiles[i].c_str(), std::ios::binary));
    if (!antStream.back()->good()) cerr << "Problem with file " << antFiles[i] << " - does it exist?" << endl;
  }

  try {
    FxKernel fxkernel(numantennas, numchannels, numffts, nbit, lo, bandwidth, iscomplex);
    fxkernel.setInputData(inputdata);
    if (readdata(subintbytes, antStream, inputdata)) return 1;
    fxkernel.setDelays(delays, antfileoffsets);

    auto starttime = std::chrono::high_resolution_clock::now();
    std::time_t time_now_t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    string starttimestring = std::ctime(&time_now_t);
    starttimestring.pop_back();
    fxkernel.process();
    auto diff = std::chrono::high_resolution_clock::now() - starttime;
    auto t1 = std::chrono::duration_cast<std::chrono::milliseconds>(diff);
    cout << "Run time was " << t1.count() << " milliseconds" << endl;
    fxkernel.saveVisibilities("vis.out");
    if (argc == 3) fxkernel.saveVisibilitiesBinary(argv[2]);
    std::ofstream ftiming("runtime.fxkernel.log");
    ftiming << "Run time was " << t1.count() << " milliseconds" << endl;
    ftiming << "Start time was " << starttimestring << endl;
  } catch (const std::exception &e) {
    cerr << e.what() << endl;
    return 1;
  }
  for (int i = 0; i < numantennas; i++) {
    delete[] inputdata[i];
    delete[] delays[i];
    delete antStream[i];
  }
  delete[] inputdata;
  delete[] delays;
  delete[] antfileoffsets;
  return 0;
}
