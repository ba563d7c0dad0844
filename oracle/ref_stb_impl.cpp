// Compiles the reference's vendored stb_image implementation (thirdparty/stb)
// into the reference harness; the reference defines it in src/Main.cpp:45,
// which the harness does not build.
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"
