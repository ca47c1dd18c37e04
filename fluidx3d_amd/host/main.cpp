// Runs the scene in setup.cpp (reference: src/main.cpp; this build has no graphics thread).
#include "setup.hpp"

int main(int argc, char** argv) {
	(void)argc; (void)argv;
	main_setup();
	return 0;
}
