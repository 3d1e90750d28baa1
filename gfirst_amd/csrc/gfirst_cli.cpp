// gfirst_cli.cpp — drop-in for the reference executable (GFIRST.cu:577-621):
//   gfirst image_volume.raw                                   -> proj_ref.raw
//   gfirst sinogram.raw normalization.raw                      doubles OSEM
//   gfirst sinogramD.raw sinogramT.raw normalization.raw       doubles + triples
//   gfirst sinogramD.raw sinogramT.raw normalization.raw VLORs.raw
// Outputs (current directory, as the reference): proj_ref.raw, then
// img_GPU_D_ITER<n>.raw / img_GPU_T_ITER<n>.raw every 20 iterations and after
// the last one, and SINOG_VLOR.raw after each iteration's triples pass.
//
// Optional leading flags (the reference compiles these in; defaults are its
// values): --geometry=ref|c0|c1 --res= --nzs= --nrad= --nang= --span= --nseg=
// --pitch= --diam= --fov= --niter= --nsub= --psf-fw= --psf-bw=
// --storage=f32|f16|bf16 --device=
//
// Host code only: it reaches the GPU exclusively through the gfr_* C ABI.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "gfirst.h"

namespace {

void usage() {
  // GFIRST.cu:589-592
  printf("1 Input  --> Projection [Input = 3D Image] \n");
  printf("2 Inputs --> Reconstruction [Input = Sinog_Doubles + Sinog_Normalization]\n");
  printf("3 Inputs --> Reconstruction Doubles & Triples [Input = Sinog_Doubles + Sinog_Triples + "
         "Sinog_Normalization] \n");
  printf("4 Inputs --> VLOR reconstruction [Input = Sinog_Doubles + Sinog_Triples + "
         "Sinog_Normalization + Vlor_file] \n");
}

bool write_raw(const char* name, const float* v, size_t n) {
  FILE* f = fopen(name, "wb");
  if (!f) return false;
  const bool ok = fwrite(v, sizeof(float), n, f) == n;
  fclose(f);
  return ok;
}

// Output image (GFIRST.cu:532-549): x*scale, zero outside radius RES/2-1.
void output_image(const gfr_geometry& g, const float* x, float scale, float* out) {
  const int res = g.res, h = res / 2;
  size_t i = 0;
  for (int z = 0; z < g.nzs; z++)
    for (int y = 0; y < res; y++)
      for (int xx = 0; xx < res; xx++, i++) {
        const float dist = (float)((xx - h) * (xx - h) + (y - h) * (y - h) - (h - 1) * (h - 1));
        out[i] = (dist > 0.f) ? 0.f : x[i] * scale;
      }
}

int die(gfr_ctx* ctx, const char* what, int rc) {
  fprintf(stderr, "gfirst: %s failed (%d): %s\n", what, rc, ctx ? gfr_last_error(ctx) : "");
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  gfr_geometry g;
  gfr_geometry_default(&g);
  gfr_recon_params p;
  p.mode = 1;
  p.niter = 60;        // GFIRST.h:20
  p.nsub = 1;          // GFIRST.h:21 (no subsets in the reference)
  p.psf_fw = 1.5f;     // GFIRST.h:26
  p.psf_bw = 1.5f;     // GFIRST.h:27
  p.jbf_r = 4;         // GFIRST.cu:512
  p.jbf_sigma3d = 0.75f;
  p.jbf_sigmai = 0.025f;
  p.jbf_coef = 0.1f;
  p.storage = GFR_STORE_F32;
  int device = 0;
  const float scale_D = (float)(3.70e-4 * 162.0);  // GFIRST.h:29-30
  const float scale_T = (float)(3.70e-4 * 162.0);
  std::vector<std::string> pos;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) {
      pos.push_back(a);
      continue;
    }
    const size_t eq = a.find('=');
    const std::string k = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
    const std::string v = eq == std::string::npos ? "" : a.substr(eq + 1);
    if (k == "geometry") {
      if (v == "c0") {
        g = {64, 32, 64, 64, 3, 11, 1, 2.0f, 106.1f, 64.0f};
      } else if (v == "c1") {
        g = {128, 159, 128, 160, 3, 53, 1, 1.63f, 165.7758f, 99.328f};
      } else if (v != "ref") {
        fprintf(stderr, "gfirst: unknown geometry %s\n", v.c_str());
        return 2;
      }
    } else if (k == "res") g.res = atoi(v.c_str());
    else if (k == "nzs") g.nzs = atoi(v.c_str());
    else if (k == "nrad") g.nrad = atoi(v.c_str());
    else if (k == "nang") g.nang = atoi(v.c_str());
    else if (k == "span") g.span = atoi(v.c_str());
    else if (k == "nseg") g.nseg = atoi(v.c_str());
    else if (k == "pitch") g.pitch = (float)atof(v.c_str());
    else if (k == "diam") g.diam_det = (float)atof(v.c_str());
    else if (k == "fov") g.fov = (float)atof(v.c_str());
    else if (k == "niter") p.niter = atoi(v.c_str());
    else if (k == "nsub") p.nsub = atoi(v.c_str());
    else if (k == "psf-fw") p.psf_fw = (float)atof(v.c_str());
    else if (k == "psf-bw") p.psf_bw = (float)atof(v.c_str());
    else if (k == "device") device = atoi(v.c_str());
    else if (k == "storage") p.storage = v == "f16" ? GFR_STORE_F16 : v == "bf16" ? GFR_STORE_BF16 : GFR_STORE_F32;
    else {
      fprintf(stderr, "gfirst: unknown option --%s\n", k.c_str());
      return 2;
    }
  }
  // argc -> mode (GFIRST.cu:588-617).
  const int nargs = (int)pos.size();
  if (nargs == 0 || nargs > 4) {
    usage();
    return 0;
  }
  const int par = nargs - 1;
  gfr_geometry_info info;
  if (gfr_geometry_check(&g, &info) != GFR_OK) {
    fprintf(stderr, "gfirst: invalid geometry\n");
    return 2;
  }
  const size_t nd = (size_t)info.ndata, nv = (size_t)info.nvox;
  gfr_ctx* ctx = nullptr;
  int rc = gfr_ctx_create(device, 0, &ctx);
  if (rc != GFR_OK) return die(nullptr, "gfr_ctx_create", rc);
  printf("Starting............ \n");

  if (par == 0) {
    // Forward projection of the input image (GFIRST.cu:152-188).
    printf("Reading Input Image for Projection ... \n");
    std::vector<float> img(nv, 0.f), sino(nd, 0.f);
    FILE* f = fopen(pos[0].c_str(), "rb");
    if (!f) {
      fprintf(stderr, "gfirst: cannot open %s\n", pos[0].c_str());
      gfr_ctx_destroy(ctx);
      return 1;
    }
    const size_t got = fread(img.data(), sizeof(float), nv, f);
    fclose(f);
    if (got != nv) fprintf(stderr, "gfirst: warning: %s holds %zu of %zu voxels\n", pos[0].c_str(), got, nv);
    void *d_img = nullptr, *d_sino = nullptr;
    if ((rc = gfr_alloc(ctx, nv * 4, &d_img)) || (rc = gfr_alloc(ctx, nd * 4, &d_sino)))
      return die(ctx, "gfr_alloc", rc);
    if ((rc = gfr_memcpy_h2d(ctx, d_img, img.data(), nv * 4, nullptr)) ||
        (rc = gfr_zero(ctx, d_sino, nd * 4, nullptr)) ||
        (rc = gfr_forward_project(ctx, &g, (float*)d_img, (float*)d_sino, 0, 1, nullptr)) ||
        (rc = gfr_memcpy_d2h(ctx, sino.data(), d_sino, nd * 4, nullptr)))
      return die(ctx, "forward projection", rc);
    if (!write_raw("proj_ref.raw", sino.data(), nd)) fprintf(stderr, "gfirst: cannot write proj_ref.raw\n");
    printf("Finish After Projection \n");
    gfr_ctx_destroy(ctx);
    return 0;
  }

  p.mode = par;
  const char* fD = pos[0].c_str();
  const char* fT = par >= 2 ? pos[1].c_str() : nullptr;
  const char* fN = pos[par == 1 ? 1 : 2].c_str();
  const char* fV = par == 3 ? pos[3].c_str() : nullptr;
  printf("\n");
  if (par == 1) printf("Mode = 1 --> Reconstruction of Doubles Sinogram \n");
  if (par == 2) printf("Mode = 2 --> Reconstruction of Doubles and Triples Sinogram \n");
  if (par == 3) printf("Mode = 3 --> Reconstruction of Doubles and Triples VLORs\n");
  printf("\nLoading Data \n");
  int isf = 0, trunc = 0;
  std::vector<float> norm2(2 * nd), d2(2 * nd), t2;
  rc = gfr_read_sinogram_file(fN, (long long)nd, norm2.data(), &isf, &trunc);
  if (rc < 0) return die(ctx, "reading normalization", rc);
  if (rc == 1) printf("Sinogram Not Found!\n");
  if (trunc) fprintf(stderr, "gfirst: warning: %s is longer than 2*NDATA words; extra words ignored\n", fN);
  printf("Normalization Sinogram loaded \n");
  rc = gfr_read_sinogram_file(fD, (long long)nd, d2.data(), &isf, &trunc);
  if (rc < 0) return die(ctx, "reading doubles", rc);
  if (rc == 1) printf("Sinogram Not Found!\n");
  if (trunc) fprintf(stderr, "gfirst: warning: %s is longer than 2*NDATA words; extra words ignored\n", fD);
  if (par >= 2) {
    t2.assign(2 * nd, 0.f);
    rc = gfr_read_sinogram_file(fT, (long long)nd, t2.data(), &isf, &trunc);
    if (rc < 0) return die(ctx, "reading triples", rc);
    if (rc == 1) printf("Sinogram Not Found!\n");
  }
  std::vector<int32_t> rec;
  long long nrec = 0;
  if (par == 3) {
    printf("Reading VLOR file \n");
    nrec = gfr_read_vlor_file(fV, nullptr, 0);
    if (nrec < 0) return die(ctx, "reading VLOR file", (int)nrec);
    rec.assign(2 * (size_t)nrec + 2, 0);
    if (nrec > 0 && gfr_read_vlor_file(fV, rec.data(), nrec) != nrec)
      return die(ctx, "reading VLOR file", GFR_EIO);
  }
  {
    // Statistics (GFIRST.cu:267-277): float sums of the RES-scaled data / RES.
    float sumd = 0.f, sumdbg = 0.f;
    for (size_t i = 0; i < nd; i++) {
      sumd += (float)g.res * d2[i];
      sumdbg += (float)g.res * d2[nd + i];
    }
    printf("  Doubles Coincidences: Prompts= %.1f Million, Background= %.1f Million\n",
           (sumd / g.res) / 1.0e6, (sumdbg / g.res) / 1.0e6);
    if (par > 1) {
      float sumt = 0.f, sumtbg = 0.f;
      for (size_t i = 0; i < nd; i++) {
        sumt += (float)g.res * t2[i];
        sumtbg += (float)g.res * t2[nd + i];
      }
      printf("  Triples Coincidences: Prompts = %.1f Million, Background= %.1f Million\n",
             (sumt / g.res) / 1.0e6, (sumtbg / g.res) / 1.0e6);
    }
  }
  gfr_recon* r = nullptr;
  rc = gfr_recon_create(ctx, &g, &p, d2.data(), par >= 2 ? t2.data() : nullptr, norm2.data(),
                        par == 3 ? rec.data() : nullptr, nrec, &r);
  if (rc != GFR_OK) return die(ctx, "gfr_recon_create", rc);
  std::vector<float> buf(nd > nv ? nd : nv), out(nv);
  if ((rc = gfr_recon_get(r, GFR_OUT_PROJ_BLANK, buf.data())) != GFR_OK) return die(ctx, "proj_ref", rc);
  write_raw("proj_ref.raw", buf.data(), nd);
  printf("\nStarting Iterations \n");
  const int ntypes = par >= 2 ? 2 : 1;
  for (int iter = 1; iter <= p.niter; iter++) {
    printf("ITER= %d  \n", iter);
    for (int ty = 0; ty < ntypes; ty++) {
      printf(ty == 0 ? "  Doubles Reconstruction \n" : "  Triples Reconstruction \n");
      for (int s = 0; s < p.nsub; s++) {
        rc = gfr_recon_step(r);
        if (rc < 0) return die(ctx, "gfr_recon_step", rc);
      }
      if (ty == 1 && par == 3) {
        if ((rc = gfr_recon_get(r, GFR_OUT_VLOR, buf.data())) != GFR_OK) return die(ctx, "vlor", rc);
        write_raw("SINOG_VLOR.raw", buf.data(), nd);
      }
    }
    if (iter % 20 == 0 || iter == p.niter) {
      char name[255];
      if ((rc = gfr_recon_get(r, GFR_OUT_XD, buf.data())) != GFR_OK) return die(ctx, "xD", rc);
      output_image(g, buf.data(), scale_D, out.data());
      snprintf(name, sizeof(name), "img_GPU_D_ITER%d.raw", iter);
      write_raw(name, out.data(), nv);
      if (par > 1) {
        if ((rc = gfr_recon_get(r, GFR_OUT_XT, buf.data())) != GFR_OK) return die(ctx, "xT", rc);
        output_image(g, buf.data(), scale_T, out.data());
        snprintf(name, sizeof(name), "img_GPU_T_ITER%d.raw", iter);
        write_raw(name, out.data(), nv);
      }
    }
  }
  gfr_recon_destroy(r);
  gfr_ctx_destroy(ctx);
  return 0;
}
