// r123_emit.cpp -- prints the Random123 OpenCL text that the reference's JIT
// puts in front of every kernel drawing random numbers (the Random123 class,
// Random123/Random123.cpp:109-135, with Philox and Threefry at 32 and 64 bits,
// vector lengths 2 and 4).  Linked by `make -C oracle ref` against the
// reference's own Random123/*.cpp and misc/PrecType.cpp, compiled where they
// lie; the text goes to oracle/_ref/r123_emitted.cl.
// TEST INFRASTRUCTURE: the emitted text is compiled for the host by
// r123_host.cl so that tests/ can compare the oracle's RNG stream with the
// reference's own generator.  This file contains no reference code.
#include <iostream>
#include "Random123.hpp"

int main()
{
    chai_internal::Random123 gen;
    gen.prec32();
    gen.prec64();
    gen.vecLength2();
    gen.vecLength4();
    gen.philox();
    gen.threefry();
    gen.print(std::cout);
    return 0;
}
