// Compiles the reference's vendored stb_image and stb_image_write
// implementations (thirdparty/stb) into the reference harness; the reference
// defines them in src/Main.cpp:45-46, which the harness does not build.
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"
#define STB_IMAGE_WRITE_IMPLEMENTATION
#include "stb_image_write.h"
