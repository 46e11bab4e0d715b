// config_maker: the reference's initial-configuration program
// (src/config_maker.cpp:518-533) over the engine.  Reads DIR/config_param.txt,
// generates the beads on the GPU and writes DIR/init_config.xyz.
//
//   config_maker DIR [SEED]      (SEED defaults to time(NULL), as the reference)
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mdpd_facade.hpp"

int main(int argc, char* argv[]) {
  if (argc != 2 && argc != 3) {
    std::fprintf(stderr, "argv[1] is target directory name.\n");
    return 1;
  }
  const std::string cdir = argv[1];
  mdpd::ConfigMaker cmaker(cdir);
  if (argc == 3) cmaker.Initialize((uint32_t)std::strtoul(argv[2], nullptr, 10));
  else cmaker.Initialize();
  cmaker.LoadParam();
  cmaker.GenParticles();
  cmaker.DumpParticleConfig();
  std::printf("Initial configuration is generated at %s\n", argv[1]);
  cmaker.ShowSystemParameter();
  return 0;
}
