// Prints the object layout of the drop-in layout classes: sizes, the
// offsets of DrawingBase's data members and the vtable slot of run().
// Compiled once against the reference's headers and once against this
// library's; the two outputs must agree for object code built against the
// reference to work with this library.  Uses no definitions (nothing is
// constructed or called), so it links without either implementation.
#include <cstdio>
#include <cstring>

#include "PHDE.h"
#include "ParHDE.h"
#include "PivotMDS.h"
#include "WParHDE.h"

template <class M>
static long member_offset(M mp) {
  long v = 0;
  std::memcpy(&v, &mp, sizeof(v));
  return v;
}

// Itanium C++ ABI: a pointer to a virtual member function holds 1 + its
// byte offset in the vtable.
template <class F>
static long vtable_slot(F f) {
  long v[2];
  std::memcpy(v, &f, sizeof(v));
  return (v[0] & 1) ? (v[0] - 1) / (long)sizeof(void *) : -1;
}

struct Probe : ParHDE {
  static void dump() {
    std::printf("rowPtrs %ld\n", member_offset(&Probe::rowPtrs));
    std::printf("adj %ld\n", member_offset(&Probe::adj));
    std::printf("weights %ld\n", member_offset(&Probe::weights));
    std::printf("n %ld\n", member_offset(&Probe::n));
    std::printf("m %ld\n", member_offset(&Probe::m));
    std::printf("numSources %ld\n", member_offset(&Probe::numSources));
    std::printf("numCoord %ld\n", member_offset(&Probe::numCoord));
    std::printf("parComp %ld\n", member_offset(&Probe::parComp));
    std::printf("eigenVectors %ld\n", member_offset(&Probe::eigenVectors));
    std::printf("coordinates %ld\n", member_offset(&Probe::coordinates));
    std::printf("ALIGN %ld\n", member_offset(&Probe::ALIGN));
    std::printf("doRand %ld\n", member_offset(&Probe::doRand));
  }
};

int main() {
  std::printf("sizeof DrawingBase %zu\n", sizeof(DrawingBase));
  std::printf("sizeof ParHDE %zu\n", sizeof(ParHDE));
  std::printf("sizeof WParHDE %zu\n", sizeof(WParHDE));
  std::printf("sizeof PHDE %zu\n", sizeof(PHDE));
  std::printf("sizeof PivotMDS %zu\n", sizeof(PivotMDS));
  std::printf("sizeof ParallelComponents %zu\n", sizeof(ParallelComponents));
  std::printf("sizeof Compare %zu\n", sizeof(Compare));
  std::printf("run slot DrawingBase %ld\n", vtable_slot(&DrawingBase::run));
  std::printf("run slot ParHDE %ld\n", vtable_slot(&ParHDE::run));
  std::printf("run slot WParHDE %ld\n", vtable_slot(&WParHDE::run));
  std::printf("run slot PHDE %ld\n", vtable_slot(&PHDE::run));
  std::printf("run slot PivotMDS %ld\n", vtable_slot(&PivotMDS::run));
  Probe::dump();
  return 0;
}
