#!/bin/sh
# Runs the generated program (built by build.sh).
exec ./main "$@"
