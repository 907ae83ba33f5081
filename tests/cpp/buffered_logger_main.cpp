// The reference's BufferedLogger test (tests/cubismup3d/TestBufferedLogger.cpp), against the
// host layer's BufferedLogger: nothing reaches the file before AUTO_FLUSH_COUNT + 1
// requests; then the first AUTO_FLUSH_COUNT lines are there, in order, and no more.
#include <cerrno>
#include <cstdio>

#include "simulation.h"

using cubismup3d::BufferedLogger;

#define CHECK(cond, ...)              \
  do {                                \
    if (!(cond)) {                    \
      std::fprintf(stderr, __VA_ARGS__); \
      return 1;                       \
    }                                 \
  } while (0)

int main()
{
  BufferedLogger logger;
  const std::string tmp = "_BufferedLogger.dat";
  const int error = std::remove(tmp.c_str());
  CHECK(error == 0 || errno == ENOENT, "remove failed: %d\n", error);
  for (int i = 1; i <= BufferedLogger::AUTO_FLUSH_COUNT; ++i) logger.get_stream(tmp) << i << '\n';
  FILE* f = std::fopen(tmp.c_str(), "r");
  CHECK(f == nullptr, "file should not be there yet\n");
  logger.get_stream(tmp) << BufferedLogger::AUTO_FLUSH_COUNT + 1 << '\n';
  f = std::fopen(tmp.c_str(), "r");
  CHECK(f != nullptr, "file should be there now\n");
  for (int i = 1; i <= BufferedLogger::AUTO_FLUSH_COUNT; ++i) {
    int x = -1;
    CHECK(std::fscanf(f, "%d ", &x) == 1 && x == i, "got %d instead of %d\n", x, i);
  }
  CHECK(std::feof(f), "file holds more than the flushed lines\n");
  std::fclose(f);
  std::puts("ok");
  return 0;
}
