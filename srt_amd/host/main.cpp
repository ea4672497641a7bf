// srtracer: command-line renderer (src/main/main.cu:44-73).
//   srtracer -c config.json [-r render.json] [-p passes.json] [-s scene.json]
// Each override file replaces the whole section of the config.
extern "C" int srt_host_run(int argc, char** argv);

int main(int argc, char** argv) { return srt_host_run(argc, argv); }
