// Regenerates the inputs of the reference unit test dmrg/Engine/tests/test_BatchedGemm.cpp:181-232
// (one patch, one operator, A 2x2, B 3x3, vin 6, std::mt19937_64(123456789), U(-1,1) drawn in the
// same order) and its expected output computed by explicit loops, as JSON on stdout.
// Build: g++ -O0 -std=c++17 gen_batchedgemm_fixture.cpp -o /tmp/gen && /tmp/gen > batchedgemm_fixture.json
#include <cstdio>
#include <random>

int main()
{
	std::mt19937_64                        rng(123456789);
	std::uniform_real_distribution<double> dist(-1.0, 1.0);
	double                                 A[2][2], B[3][3], vin[6];
	for (int i = 0; i < 2; i++)
		for (int j = 0; j < 2; j++)
			A[i][j] = dist(rng);
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++)
			B[i][j] = dist(rng);
	for (int i = 0; i < 6; i++)
		vin[i] = dist(rng);
	double X[3][2], BX[3][2], Y[3][2];
	for (int col = 0; col < 2; col++)
		for (int row = 0; row < 3; row++)
			X[row][col] = vin[row + col * 3];
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 2; j++) {
			double s = 0;
			for (int k = 0; k < 3; k++)
				s += B[i][k] * X[k][j];
			BX[i][j] = s;
		}
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 2; j++) {
			double s = 0;
			for (int k = 0; k < 2; k++)
				s += BX[i][k] * A[j][k];
			Y[i][j] = s;
		}
	// Matrices are emitted column major, as PsimagLite::Matrix stores them.
	printf("{\n \"source\": \"dmrg/Engine/tests/test_BatchedGemm.cpp:181-232\",\n");
	printf(" \"L\": [2], \"R\": [3], \"nC\": 1, \"np\": 1,\n \"A\": [");
	for (int j = 0; j < 2; j++)
		for (int i = 0; i < 2; i++)
			printf("%s%.17g", (i + j) ? ", " : "", A[i][j]);
	printf("],\n \"B\": [");
	for (int j = 0; j < 3; j++)
		for (int i = 0; i < 3; i++)
			printf("%s%.17g", (i + j) ? ", " : "", B[i][j]);
	printf("],\n \"vin\": [");
	for (int i = 0; i < 6; i++)
		printf("%s%.17g", i ? ", " : "", vin[i]);
	printf("],\n \"expected\": [");
	for (int col = 0; col < 2; col++)
		for (int row = 0; row < 3; row++)
			printf("%s%.17g", (row + col) ? ", " : "", Y[row][col]);
	printf("],\n \"tolerance\": 1e-12\n}\n");
	return 0;
}
