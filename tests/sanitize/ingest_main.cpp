// Reads each file named on the command line with the scene-ingest readers
// (include/ocr_obj.h): *.png through ocr_png_load, anything else through
// ocr_obj_load + ocr_obj_get.  Prints one status line per file; built with
// the sanitizers, a bad access in a reader aborts the run.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/ocr_obj.h"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; i++) {
    const char* path = argv[i];
    const size_t n = strlen(path);
    if (n > 4 && !strcmp(path + n - 4, ".png")) {
      uint32_t w = 0, h = 0;
      uint8_t* px = nullptr;
      const int rc = ocr_png_load(path, &w, &h, &px);
      unsigned sum = 0;
      if (rc == 0)
        for (size_t k = 0; k < (size_t)w * h * 4; k++) sum += px[k];
      printf("%s png %d %ux%u %u\n", path, rc, w, h, sum);
      ocr_png_free(px);
    } else {
      ocr_obj_mesh* m = nullptr;
      const int rc = ocr_obj_load(path, 1.5f, &m);
      uint32_t nt = 0, np = 0;
      if (rc == 0) {
        nt = ocr_obj_ntri(m);
        np = ocr_obj_nparts(m);
        std::vector<ocr_triangle> tris(nt);
        std::vector<ocr_obj_part> parts(np);
        ocr_obj_get(m, tris.data(), parts.data());
        ocr_obj_free(m);
      }
      printf("%s obj %d %u %u\n", path, rc, nt, np);
    }
  }
  return 0;
}
