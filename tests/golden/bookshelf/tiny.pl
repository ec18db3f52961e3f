UCLA pl 1.0

a 0 0 : N
b 0 0 : N
c 0 0 : N
pad 80 10 : N /FIXED
