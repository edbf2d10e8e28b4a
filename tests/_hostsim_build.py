"""Builds the host-only runtime binaries of tests/hostsim with make, one
build at a time across processes (parallel test workers would otherwise
run make concurrently and start a binary another make is still writing)."""
import fcntl
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostsim")


def make(*targets):
    with open(os.path.join(HERE, ".make.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)
