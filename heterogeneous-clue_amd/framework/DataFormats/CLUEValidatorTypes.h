// Conditions products of the CLUE2D validator. OutputDirPath is the reference's product
// (src/serial/plugin-CLUEValidator/CLUEValidatorTypes.h:4-6): the output CSV path derived from the
// input file, whose directory holds reference/ref_<N>k_20_25_2.csv. ValidationMode is optional and
// has no counterpart in the reference: strict = true adds a byte-for-byte row check on top of the
// reference's tolerance check.
#ifndef CLUE_DF_CLUEValidatorTypes_h
#define CLUE_DF_CLUEValidatorTypes_h

#include <filesystem>

struct OutputDirPath {
  std::filesystem::path outFile;
};

struct ValidationMode {
  bool strict = false;
};

#endif
